"""Bundles run in parts: a bundle of more quadrature nodes than Kernel.part_nodes runs as sub-bundles of consecutive
whole groups, one visit plan at a time (the plan of a part is built for its launch and freed after it), so that the
reference's production quadrature (nmu 12, nphi 16, path_n 12) fits in device memory with int32 visit ids per part.

CPU: the group split and the sub-bundle views (rebased group, edge and ray-group offsets) against the bundle built
from the same cells alone. GPU: first light, the warm pass with a boundary, the ground pass and a whole solve of a
bundle run in parts equal the bundle run whole, bit for bit (every ray and group sees the same arithmetic)."""

from dataclasses import replace

import numpy as np
import pytest
import torch

from sosfuse2_amd import api, atmosphere, geometry
from sosfuse2_amd.kernel import DeviceBundle, check_plan_size, split_groups


def test_split_groups():
    per = np.array([5, 3, 0, 7, 2, 9, 1])
    assert split_groups(per, 100) == [(0, 7)]
    parts = split_groups(per, 8)
    assert parts == [(0, 3), (3, 4), (4, 5), (5, 6), (6, 7)]
    for g0, g1 in parts:
        assert per[g0:g1].sum() <= 8 or g1 - g0 == 1
    assert [g for p in parts for g in range(*p)] == list(range(len(per)))
    assert split_groups(per, 1) == [(0, 1), (1, 3), (3, 4), (4, 5), (5, 6), (6, 7)]  # a 0-node group joins its left
    assert split_groups(np.zeros(0, int), 4) == []
    with pytest.raises(ValueError, match='E_PART_NODES'):
        split_groups(per, 0)
    with pytest.raises(OverflowError, match='E_PLAN_SIZE'):
        split_groups([2**29 + 1], 2**27)
    assert check_plan_size(2**29 - 1) == 2**29 - 1


def test_bundle_part_views_match_the_cells_alone():
    st = atmosphere.config0_state()
    cfg = dict(nr=4, chi_grid='tiny', nmu=1, nphi=2, path_n=4, L=16, top_km=100.0, workers=1)
    grid = api.make_grid(cfg)
    packs = geometry.cell_packs(st, grid)
    whole = DeviceBundle(geometry.Bundle(packs), 'cpu')
    for g0, g1 in ((0, 1), (3, 17), (25, 40), (0, 40)):
        part = whole.part(g0, g1)
        alone = DeviceBundle(geometry.Bundle(packs[g0:g1]), 'cpu')
        assert (part.groups, part.rays, part.nq) == (alone.groups, alone.rays, alone.nq)
        for name in ('group_off', 'starts', 'targets', 'ray_group', 'dirs', 'dloc', 'weights', 'edge_off', 'edges',
                     'lower', 'end_j', 'end_f', 'end_col', 'hits'):
            assert torch.equal(getattr(part, name), getattr(alone, name)), name
        assert part.samples == alone.samples
        assert part.span[:2] == (g0, g1)
    with pytest.raises(ValueError, match='E_BUNDLE_PART'):
        whole.part(5, 5)


@pytest.mark.gpu
def test_parts_equal_the_whole_bundle(gpu):
    """part_nodes below a third of the bundle's nodes: first light, the warm pass with a boundary and the ground
    pass, then a solve, equal the single-plan runs bit for bit."""
    from tests.test_gpu_parity import physical_field

    cfg = dict(nr=6, chi_grid='tiny', nmu=1, nphi=2, path_n=4, L=16, top_km=80.0, observer_km=0.1, query_n=6,
               anderson_depth=4)
    wl = atmosphere.log_spaced_wavelengths(3, 420, 680)
    s, scope = api.create_configured(None, cfg, wl, 'hip', replace(atmosphere.config0_state(), top_km=80.0))
    k = s.kernel
    geo = scope.solver(s, log=None).geometry
    db = geo.bundle
    dc = k.constants(s, db.nq)
    assert k.parts(db, dc) == [db]
    field = torch.from_numpy(np.ascontiguousarray(
        np.concatenate([physical_field(s, 3), np.repeat(physical_field(s, 3)[..., -1:], dc.G - s.W, -1)], -1))).cuda()
    bnd = torch.rand((s.geo.nc, dc.G), device='cuda') * 1e-2

    def runs():
        m1, f1 = geo.evaluate(s, None, first=True, phase=0)
        m2, f2 = geo.evaluate(s, field, phase=0, boundary=bnd)
        _, f3 = geo.evaluate(s, field, phase=1, ground_only=True)
        x, info = scope.solver(s, log=None).solve(40, 1e-5)
        return [t.cpu() for t in (m1, f1, m2, f2, f3, x)], info['evaluations']

    whole, n_whole = runs()
    nodes = db.plan.nodes
    db.plan = None
    db._parts = None
    k.part_nodes = nodes // 3
    try:
        parts = k.parts(db, dc)
        assert len(parts) >= 3 and sum(p.rays for p in parts) == db.rays
        print(f'parts: {len(parts)} parts of at most {k.part_nodes} of {nodes} nodes', flush=True)
        split, n_split = runs()
        assert all(p.plan is None for p in parts)  # plans are freed after each launch
    finally:
        k.part_nodes = 2**27
        db._parts = None
    assert n_split == n_whole
    for a, b in zip(whole, split):
        assert torch.equal(a, b)
