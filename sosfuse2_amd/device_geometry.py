"""Grid geometry and column integrals computed on the device (the reference builds them on the host):

- `cell_bundles`: the incident-direction quadrature of every cell (`src/sosfuse2/rays/incoming_rings.py:63-126`), the
  ray packs (`src/sosfuse2/solve/ray_geometry.py:28-82`) and their compacted CSR (`:148-180`), with the end columns
  of ground-hitting rays (`:53-65`), straight into the device bundles of the warm pass and the ground pass;
- `sun_table`: the sun-column table (`src/sosfuse2/rays/shell_segments.py:230-307`);
- `end_columns`, `pool_columns`: the end columns of host packs and the path columns of the ground pools
  (`src/sosfuse2/solve/engine_core.py:103-147`).

The FP64 arithmetic restates `geometry.incoming`, `geometry.pack`, `geometry.compact_edges`,
`geometry.boundaries` and `geometry.columns` (kernels in csrc/geometry.hip)."""

import ctypes
import math

import numpy as np
import torch

from . import geometry, hip


class _ColumnSpec:
    """The sos_columns struct of a state's profile shapes for one node count, with the device arrays it points to."""

    def __init__(self, state, top_km, nq, device):
        spec = state.profiles
        u, wu = geometry.gl(nq)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a, 'f8')).to(device)
        self.u, self.w = t(u), t(wu)
        self.table = t(spec.table_array()) if spec.table is not None else None
        self.struct = hip.Columns(float(top_km), int(nq), hip.ptr(self.u), hip.ptr(self.w))
        for k in range(6):
            self.struct.profile_kind[k] = int(spec.kinds[k])
            self.struct.profile_a[k] = float(spec.a[k])
            self.struct.profile_b[k] = float(spec.b[k])
        if self.table is not None:
            self.struct.profile_table_dev = hip.ptr(self.table)
            self.struct.profile_table_n = int(self.table.shape[1])


def _device(device):
    hip.mark_started()
    return torch.device(device)


def sun_table(state, domain, scale=1, device='cuda'):
    """geometry.sun_table on the device: columns toward the sun on (81 scale + 1 heights, 400 scale + 1 x)."""
    dev = _device(device)
    nh = 80 * scale + 1
    h = np.linspace(0, math.sqrt(domain.top_km), nh) ** 2
    x = np.linspace(0, 1, 400 * scale + 1)
    cs = _ColumnSpec(state, domain.top_km, 32 * scale, dev)
    hd = torch.from_numpy(h).to(dev)
    xd = torch.from_numpy(x).to(dev)
    out = torch.empty((nh, len(x), 6), dtype=torch.float64, device=dev)
    hip.call('sos_sun_table', cs.struct, nh, len(x), hip.ptr(hd), hip.ptr(xd), hip.ptr(out),
             torch.cuda.current_stream(dev).cuda_stream)
    return geometry.SunTable(h, x, out.cpu().numpy())


def ray_columns(state, domain, starts, dirs, group=None, hit=None, n=24, device='cuda'):
    """Columns [rays, 6] from starts (per ray, or per group with group[i] the ray's group) along dirs to each ray's
    end, zero where hit is False (geometry.columns over geometry.boundaries, on the device)."""
    dev = _device(device)
    dirs = np.ascontiguousarray(dirs, 'f8')
    m = len(dirs)
    if m == 0:
        return np.zeros((0, 6))
    cs = _ColumnSpec(state, domain.top_km, n, dev)
    t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dt)).to(dev)
    sd, dd = t(starts, 'f8'), t(dirs, 'f8')
    gd = None if group is None else t(group, 'i4')
    hd = None if hit is None else t(np.asarray(hit, bool), 'u1')
    out = torch.empty((m, 6), dtype=torch.float64, device=dev)
    hip.call('sos_ray_columns', cs.struct, m, hip.ptr(sd), hip.ptr(gd), hip.ptr(dd), hip.ptr(hd), hip.ptr(out),
             torch.cuda.current_stream(dev).cuda_stream)
    return out.cpu().numpy()


def end_columns(state, domain, packs, device='cuda'):
    """Set g['end_columns'] of every pack (geometry.with_end_columns, n = 24) from one device launch."""
    counts = [len(g['d']) for g in packs]
    starts = np.array([g['p'][0] if len(g['d']) else np.zeros(3) for g in packs])
    group = np.repeat(np.arange(len(packs), dtype='i4'), counts)
    dirs = np.concatenate([g['d'] for g in packs])
    hit = np.concatenate([g['hits_ground'] for g in packs])
    col = ray_columns(state, domain, starts, dirs, group, hit, 24, device)
    off = np.r_[0, np.cumsum(counts)]
    for i, g in enumerate(packs):
        g['end_columns'] = col[off[i] : off[i + 1]]
    return packs


def pool_columns(state, domain, pools, sun, device='cuda'):
    """Fill the columns of ground pools built with deferred columns (geometry.ground_pools(columns_too=False)): the
    path from the cell to the ground point on the device plus the sun table's columns at the ground point."""
    if pools.path is None:
        return pools
    starts, dirs, ground = pools.path
    path = ray_columns(state, domain, starts, dirs, None, None, 24, device)
    sc, _ = geometry.sun_lookup(sun, ground, np.maximum(geometry.R, np.linalg.norm(ground, axis=-1)), domain.top_km)
    pools.column = path + sc
    pools.path = None
    return pools


def cell_bundles(state, grid, device='cuda'):
    """(bundle of every cell, bundle of the first height row's cells) on the device: incident quadrature, packs, CSR
    edges and end columns built by HIP kernels, the same arrays geometry.Bundle(geometry.cell_packs(...)) uploads."""
    from .kernel import DeviceBundle

    dev = _device(device)
    st = lambda: torch.cuda.current_stream(dev).cuda_stream
    t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dt)).to(dev)
    dom = grid.domain
    hd, chid = t(grid.h, 'f8'), t(grid.chi, 'f8')
    u, wu = geometry.gl(grid.nmu)
    ud, wd = t(u, 'f8'), t(wu, 'f8')
    levels = [-math.sqrt(max(0.0, (geometry.R + h) ** 2 - geometry.R ** 2)) for h in dom.ring_heights]
    inc = hip.Incident(grid.nh, grid.nc, grid.nmu, grid.nphi, len(levels), hip.ptr(hd), hip.ptr(chid),
                       (ctypes.c_double * 8)(*levels), hip.ptr(ud), hip.ptr(wd))
    per = grid.nphi + 1
    counts = torch.empty(grid.ns * per, dtype=torch.int64, device=dev)
    hip.call('sos_incident_count', inc, hip.ptr(counts), st())
    off = torch.zeros_like(counts)
    off[1:] = torch.cumsum(counts, 0)[:-1]
    cell_rays = counts.view(grid.ns, per).sum(1)
    group_off = torch.zeros(grid.ns + 1, dtype=torch.int64, device=dev)
    group_off[1:] = torch.cumsum(cell_rays, 0)
    rays = int(group_off[-1].item())
    dirs = torch.empty((max(rays, 1), 3), dtype=torch.float64, device=dev)
    weights = torch.empty(max(rays, 1), dtype=torch.float64, device=dev)
    ray_cell = torch.empty(max(rays, 1), dtype=torch.int32, device=dev)
    hip.call('sos_incident_fill', inc, hip.ptr(off), hip.ptr(dirs), hip.ptr(weights), hip.ptr(ray_cell), st())
    # cell starts and target frames (host FP64, as geometry.cell_pack: the frame at the cell position)
    ih, ic = np.divmod(np.arange(grid.ns), grid.nc)
    p, _, _ = geometry.frame(geometry.R + grid.h[ih], grid.chi[ic])
    B = geometry.basis_at(p)
    starts, targets = t(p, 'f8'), t(B.reshape(-1, 9), 'f8')
    cs = _ColumnSpec(state, dom.top_km, 24, dev)
    edge_count = torch.empty(max(rays, 1), dtype=torch.int64, device=dev)
    sampled = torch.empty(max(rays, 1), dtype=torch.int64, device=dev)
    hip.call('sos_pack_count', cs.struct, grid.nc, hip.ptr(chid), hip.ptr(starts), rays, hip.ptr(ray_cell),
             hip.ptr(dirs), hip.ptr(edge_count), hip.ptr(sampled), st())
    edge_off = torch.zeros(rays + 1, dtype=torch.int64, device=dev)
    edge_off[1:] = torch.cumsum(edge_count[:rays], 0)
    E = int(edge_off[-1].item())
    edges = torch.empty(max(E, 1), dtype=torch.float64, device=dev)
    lower = torch.empty(max(rays, 1), dtype=torch.int16, device=dev)
    end_j = torch.empty(max(rays, 1), dtype=torch.int32, device=dev)
    end_f = torch.empty(max(rays, 1), dtype=torch.float32, device=dev)
    hits = torch.empty(max(rays, 1), dtype=torch.uint8, device=dev)
    dloc = torch.empty((max(rays, 1), 3), dtype=torch.float32, device=dev)
    hip.call('sos_pack_fill', cs.struct, grid.nc, hip.ptr(chid), hip.ptr(starts), hip.ptr(targets), rays,
             hip.ptr(ray_cell), hip.ptr(dirs), hip.ptr(edge_off), hip.ptr(edges), hip.ptr(lower), hip.ptr(end_j),
             hip.ptr(end_f), hip.ptr(hits), hip.ptr(dloc), st())
    end_col = torch.empty((max(rays, 1), 6), dtype=torch.float64, device=dev)
    hip.call('sos_ray_columns', cs.struct, rays, hip.ptr(starts), hip.ptr(ray_cell), hip.ptr(dirs), hip.ptr(hits),
             hip.ptr(end_col), st())
    samples = int(sampled[:rays].sum().item()) * grid.path_n
    full = dict(group_off=group_off, starts=starts, targets=targets.float(), ray_group=ray_cell, dirs=dirs, dloc=dloc,
                weights=weights.float(), edge_off=edge_off, edges=edges, lower=lower, end_j=end_j, end_f=end_f,
                end_col=end_col.float(), hits=hits)
    bundle = DeviceBundle.from_tensors(grid.ns, rays, grid.path_n, samples, full)
    # the ground pass: the first height row's cells, a prefix of every array
    nc = grid.nc
    rg = int(group_off[nc].item())
    eg = int(edge_off[rg].item())
    ground = dict(group_off=group_off[: nc + 1], starts=starts[:nc], targets=full['targets'][:nc],
                  ray_group=ray_cell[:rg], dirs=dirs[:rg], dloc=dloc[:rg], weights=full['weights'][:rg],
                  edge_off=edge_off[: rg + 1], edges=edges[:eg], lower=lower[:rg], end_j=end_j[:rg],
                  end_f=end_f[:rg], end_col=full['end_col'][:rg], hits=hits[:rg])
    gsamples = int(sampled[:rg].sum().item()) * grid.path_n
    return bundle, DeviceBundle.from_tensors(nc, rg, grid.path_n, gsamples, ground)
