"""Host-side rules of the engine that run without a GPU: the observer monitor sets, the visit-plan index limit, the
profile-shape check of states sharing a grid, and the process-pool choice of the host preparation."""

import numpy as np
import pytest

from oracle.solve import monitor_elevations
from sosfuse2_amd import api, atmosphere, engine, geometry, kernel


def _context(top_km):
    cfg = dict(nr=4, chi_grid='tiny', nmu=1, nphi=2, path_n=4, L=16, top_km=top_km)
    grid = api.make_grid(cfg)
    optics = atmosphere.lane_optics(atmosphere.State(top_km=top_km), [550.0])
    return engine.Context(grid, optics, kernel=None, sun=object(), pools=object())


@pytest.mark.parametrize('top_km, stop, record', [
    (80.0, (2, -3, -6, -9, -12), ()),
    (100.0, (2, -3, -6, -9, -12, -15), (-18,)),
    (120.0, (2, -3, -6, -9, -12, -15), (-18,)),
])
def test_monitor_sets_follow_the_top(top_km, stop, record):
    """Above an 80 km top the stop set gains -15 deg (six x seven boxes) and -18 deg is recorded only
    (`src/sosfuse2/foundation/top.py:55-57`); the oracle states the same rule on its own."""
    s = _context(top_km)
    assert s.monitors == tuple(float(v) for v in stop)
    assert s.record_monitors == tuple(float(v) for v in record)
    o_stop, o_record = monitor_elevations(top_km)
    assert tuple(o_stop) == s.monitors and tuple(o_record) == s.record_monitors


def test_plan_size_limit_raises_before_the_ids_wrap():
    assert kernel.check_plan_size(1000) == 1000
    assert kernel.check_plan_size(kernel.INDEX_LIMIT // 4) == kernel.INDEX_LIMIT // 4
    with pytest.raises(OverflowError, match='E_PLAN_SIZE'):
        kernel.check_plan_size(kernel.INDEX_LIMIT // 4 + 1)


def test_states_of_other_profile_shapes_are_rejected():
    base = atmosphere.config1_state()
    atmosphere.check_profile_shapes(atmosphere.with_albedo(base, 0.3), base)  # optical amounts may differ
    from dataclasses import replace

    for key, value in (('aerosol_height_km', 2.0), ('top_km', 100.0), ('ozone_center_km', 25.0)):
        with pytest.raises(ValueError, match='E_BATCH_PROFILE'):
            atmosphere.check_profile_shapes(replace(base, **{key: value}), base)
    a = atmosphere.lane_optics(base, [500.0])
    b = atmosphere.lane_optics(replace(base, aerosol_height_km=2.0), [600.0])
    with pytest.raises(ValueError, match='E_BATCH_PROFILE'):
        atmosphere.concat_optics([a, b])


def test_scope_rejects_a_state_of_other_profile_shapes():
    from types import SimpleNamespace
    from dataclasses import replace

    base = atmosphere.config0_state()
    scope = SimpleNamespace(state=base)
    with pytest.raises(ValueError, match='E_BATCH_PROFILE'):
        api.create_configured(None, {}, [550.0], 'hip', replace(base, aerosol_height_km=2.0), scope=scope)


def test_host_preparation_uses_threads_once_the_gpu_runtime_is_up(monkeypatch):
    """A process pool forked from a process that holds the GPU runtime is not started: the preparation switches to
    threads, with the same packs as the serial path."""
    cfg = dict(nr=8, chi_grid='tiny', nmu=1, nphi=2, path_n=4, L=16, top_km=100.0)
    grid = api.make_grid(cfg)
    st = atmosphere.config0_state()
    monkeypatch.setattr(geometry, 'gpu_runtime_started', lambda: True)
    threaded = geometry.cell_packs(st, grid, workers=4)
    serial = geometry.cell_packs(st, grid, workers=1)
    assert len(threaded) == len(serial) == grid.ns
    for a, b in zip(threaded, serial):
        np.testing.assert_array_equal(a['d'], b['d'])
        np.testing.assert_array_equal(a['end_columns'], b['end_columns'])
