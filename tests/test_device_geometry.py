"""Column integrals on the device (sos_sun_table, sos_ray_columns) against the host FP64 restatement of the same
geometry (geometry.sun_table, geometry.with_end_columns, geometry.ground_pools): the sun-column table, the end
columns of ground-hitting rays and the ground pools, for the configs' analytic profiles (with a sulfate layer, top
above 80 km) and for the reference's legacy shapes (triangle ozone, tabulated NO2, top 80 km). Both evaluate the
same quadrature in FP64 with different summation orders: agreement to 1e-12 of each column's largest value."""

import copy
from types import SimpleNamespace

import numpy as np
import pytest

from sosfuse2_amd import api, atmosphere, geometry

pytestmark = pytest.mark.gpu


def _states():
    from tests.test_seam import legacy_state

    return [(atmosphere.config1_state(), 120.0), (legacy_state()[0], 80.0)]


def close(a, b, tol=1e-12):
    a, b = np.asarray(a), np.asarray(b)
    scale = np.maximum(np.abs(b).max(axis=tuple(range(b.ndim - 1)), keepdims=True), 1e-300)
    return float(np.max(np.abs(a - b) / scale))


@pytest.mark.parametrize('which', [0, 1])
def test_sun_table_on_the_device(gpu, which):
    from sosfuse2_amd import device_geometry as dg

    st, top = _states()[which]
    dom = geometry.Domain(top)
    host = geometry.sun_table(st, dom, 1)
    dev = dg.sun_table(st, dom, 1)
    np.testing.assert_array_equal(dev.h, host.h)
    np.testing.assert_array_equal(dev.x, host.x)
    assert close(dev.col, host.col) < 1e-12


def test_sun_table_scale_4_on_the_device(gpu):
    """The reference's production sun table scale (4: 321 heights x 1601 cosines, 128 nodes per segment) against
    host rows at a few heights."""
    from sosfuse2_amd import device_geometry as dg

    st, top = _states()[1]
    dom = geometry.Domain(top)
    dev = dg.sun_table(st, dom, 4)
    assert dev.col.shape == (321, 1601, 6)
    rows = [0, 1, 97, 200, 320]
    host = geometry._sun_rows((st, dom, dev.h[rows], dev.x, 4))
    assert close(dev.col[rows], host) < 1e-12


@pytest.mark.parametrize('which', [0, 1])
def test_end_and_pool_columns_on_the_device(gpu, which):
    from sosfuse2_amd import device_geometry as dg

    st, top = _states()[which]
    cfg = dict(nr=6, chi_grid='tiny', nmu=2, nphi=4, path_n=4, L=16, top_km=top, workers=1)
    grid = api.make_grid(cfg)
    sun = geometry.sun_table(st, grid.domain, 1)
    host_packs = geometry.cell_packs(st, grid)
    dev_packs = dg.end_columns(st, grid.domain, geometry.cell_packs(st, grid, end_columns=False))
    hc = np.concatenate([g['end_columns'] for g in host_packs])
    dc = np.concatenate([g['end_columns'] for g in dev_packs])
    hits = np.concatenate([g['hits_ground'] for g in host_packs])
    assert hits.sum() > 50 and np.all(dc[~hits] == 0)
    assert close(dc, hc) < 1e-12
    hp = geometry.ground_pools(st, grid, sun)
    dp = dg.pool_columns(st, grid.domain, geometry.ground_pools(st, grid, None, columns_too=False), sun)
    np.testing.assert_array_equal(dp.off, hp.off)
    np.testing.assert_array_equal(dp.loc, hp.loc)
    np.testing.assert_array_equal(dp.factor, hp.factor)
    assert len(hp.column) > 100 and close(dp.column, hp.column) < 1e-12


def sampled_intervals(edge_off, edges, lower):
    """(ray, start, end, lower flag) of every compacted interval wider than 1e-8 km."""
    edge_off = np.asarray(edge_off, 'i8')
    bits = np.asarray(lower).astype('u2').astype('i8')
    n = np.diff(edge_off) - 1  # intervals per ray
    ray = np.repeat(np.arange(len(n)), np.maximum(n, 0))
    first = np.repeat(edge_off[:-1], np.maximum(n, 0))
    rank = np.arange(len(ray)) - np.repeat(np.cumsum(np.maximum(n, 0)) - np.maximum(n, 0), np.maximum(n, 0))
    a, e = edges[first + rank], edges[first + rank + 1]
    flag = (bits[ray] >> rank) & 1
    keep = e - a > 1e-8
    return ray[keep], a[keep], e[keep], flag[keep]


@pytest.mark.parametrize('which', [0, 1])
def test_cell_bundles_on_the_device(gpu, which):
    """device_geometry.cell_bundles (incident quadrature, packs, CSR, end columns by HIP kernels) against
    geometry.Bundle of the host packs: the same rays in the same order, equal integer arrays, directions and edges
    to 1e-12 (the device and host transcendental functions may differ in the last bit, which the sinh grading
    near the horizon amplifies)."""
    from sosfuse2_amd import device_geometry as dg

    st, top = _states()[which]
    cfg = dict(nr=7, chi_grid='tiny', nmu=2, nphi=4, path_n=4, L=16, top_km=top, workers=1)
    grid = api.make_grid(cfg)
    bundle, ground = dg.cell_bundles(st, grid)
    for b, packs in ((bundle, geometry.cell_packs(st, grid)), (ground, geometry.cell_packs(st, grid)[: grid.nc])):
        host = geometry.Bundle(packs)
        assert (b.groups, b.rays, b.samples) == (host.groups, host.rays, host.samples)
        get = lambda name: getattr(b, name).cpu().numpy()
        for name in ('group_off', 'hits', 'ray_group'):
            np.testing.assert_array_equal(get(name), getattr(host, name).astype(get(name).dtype), err_msg=name)
        np.testing.assert_allclose(get('dirs'), host.dirs, rtol=0, atol=1e-12)
        # the sampled intervals (wider than 1e-8 km) with their lower-end flags; intervals narrower than that carry no
        # nodes and may come and go with the last bit of a boundary at an interior shell where a ray starts
        sd = sampled_intervals(get('edge_off'), get('edges'), get('lower'))
        sh = sampled_intervals(host.edge_off, host.edges, host.lower)
        np.testing.assert_array_equal(sd[0], sh[0])
        np.testing.assert_array_equal(sd[3], sh[3])
        np.testing.assert_allclose(sd[1], sh[1], rtol=1e-12, atol=1e-9)
        np.testing.assert_allclose(sd[2], sh[2], rtol=1e-12, atol=1e-9)
        np.testing.assert_allclose(get('starts'), host.starts, rtol=1e-15, atol=0)
        np.testing.assert_allclose(get('weights'), host.weights, rtol=1e-6, atol=0)
        np.testing.assert_allclose(get('dloc'), host.dloc, rtol=0, atol=2e-7)
        # a ray ending on a chi node may take the cell on either side (index j with fraction 1, or j + 1 with 0): the
        # interpolation position j + f is what must agree
        np.testing.assert_allclose(get('end_j') + get('end_f').astype('f8'), host.end_j + host.end_f.astype('f8'),
                                   rtol=0, atol=2e-6)
        assert np.mean(get('end_j') == host.end_j) > 0.99
        np.testing.assert_allclose(get('end_col'), host.end_col, rtol=2e-6, atol=1e-30)


def test_configured_scope_uses_device_geometry(gpu):
    """create_configured builds its bundles, sun table and ground pools on the device; they equal a host-built
    scope's, and the device-built geometry's solve matches the host-built one's."""
    cfg = dict(nr=5, chi_grid='tiny', nmu=1, nphi=2, path_n=4, L=16, top_km=100.0, query_n=6, workers=1)
    st = atmosphere.config0_state()
    s, scope = api.create_configured(None, cfg, [550.0], 'hip', st)
    assert scope.device_geometry and scope.bundles is not None and scope.pools.path is None
    host = api.Scope(dict(cfg, sun_scale=1), st, SimpleNamespace(L=16))
    assert close(scope.sun.col, host.sun.col) < 1e-12
    assert close(scope.pools.column, host.pools.column) < 1e-12
    x_dev, i_dev = scope.solver(s, log=None).solve(40, 1e-5)
    from sosfuse2_amd.solver import Geometry, Solver

    sv = Solver(s, Geometry(s, host.packs, s.kernel), depth=4, log=None)
    x_host, i_host = sv.solve(40, 1e-5)
    assert i_dev['evaluations'] == i_host['evaluations']
    err = float((x_dev - x_host).abs().max() / x_host.abs().max())
    assert err < 1e-6, err
