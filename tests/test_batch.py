"""Batched workload: the seeded state distribution and the lane packing (host logic, no GPU)."""

import numpy as np

from sosfuse2_amd import batch


def test_seeded_items_follow_the_batched_distribution():
    items = batch.items_from_seed(64, seed=3)
    assert len(items) == 64
    aod = np.array([it.state.aod550 for it in items])
    g = np.array([it.state.aerosol_g for it in items])
    albedo = np.array([it.state.albedo for it in items])
    nw = np.array([len(it.wavelengths) for it in items])
    ne = np.array([len(it.elevations) for it in items])
    assert aod.min() >= 0.02 and aod.max() <= 0.5
    assert g.min() >= 0.6 and g.max() <= 0.8
    assert albedo.min() >= 0.0 and albedo.max() <= 0.3
    assert nw.min() >= 1 and nw.max() <= 30 and len(set(nw)) > 10
    assert ne.min() >= 2 and ne.max() <= 25
    for it in items:
        assert it.elevations[0] == 6.0 and all(a - b == 1.0 for a, b in zip(it.elevations, it.elevations[1:]))
    again = batch.items_from_seed(64, seed=3)
    assert [it.state.aod550 for it in again] == list(aod)


def test_pack_fills_lanes_without_splitting_states():
    items = batch.items_from_seed(64, seed=3)
    chunks = batch.pack(items, lanes=32, max_states=4)
    seen = sorted(i for ch in chunks for i in ch.members)
    assert seen == list(range(64))
    for ch in chunks:
        assert ch.lanes == sum(len(items[i].wavelengths) for i in ch.members) <= 32
        assert 1 <= len(ch.members) <= 4
    total = sum(len(it.wavelengths) for it in items)
    # first-fit decreasing stays within a few chunks of the lane bound
    assert len(chunks) <= -(-total // 32) + 4
