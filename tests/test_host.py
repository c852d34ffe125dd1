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


def test_table_npz_has_the_reference_layout(tmp_path):
    """api.save_table writes the reference's table.npz keys, dtypes and axis order; read_table (the reference's
    checks) reads it back; Stokes Q and U ride along as extra arrays."""
    wl, hs, el, az = api.axes(api.CONFIG0)
    rng = np.random.default_rng(0)
    table = rng.uniform(1e-6, 1e-3, size=(len(hs), len(el), len(az), len(wl), 3)).astype('f4')
    path = api.save_table(tmp_path / 'table.npz', table, wl, hs, el, az)
    with np.load(path, allow_pickle=False) as f:
        assert set(f.files) == {'radiance', 'wavelength_nm', 'solar_elevation_deg', 'view_elevation_deg',
                                'relative_azimuth_deg', 'units', 'stokes_q', 'stokes_u'}
        assert f['radiance'].dtype == np.float32 and f['radiance'].shape == (2, 4, 6, 1)
        assert all(f[k].dtype == np.float32 for k in api.TABLE_AXES)
        assert f['units'].item() == 'W m^-2 sr^-1 nm^-1'
        np.testing.assert_array_equal(f['stokes_q'], table[..., 1])
    axes, radiance = api.read_table(path)
    np.testing.assert_array_equal(radiance, table[..., 0])
    np.testing.assert_array_equal(axes[1], np.float32(el))
    np.testing.assert_array_equal(axes[0], [-3, -6])
    np.testing.assert_array_equal(axes[1], [78.75, 56.25, 33.75, 11.25])  # the reference's reduced view rows
    np.testing.assert_array_equal(axes[2], [15, 45, 75, 105, 135, 165])
    with pytest.raises(ValueError, match='E_TABLE_SHAPE'):
        api.save_table(tmp_path / 'bad.npz', table[:, :2], wl, hs, el, az)
    assert not (tmp_path / 'table.npz.tmp').exists()


def test_state_queue_packs_prepares_and_fails_per_chunk(tmp_path):
    """The queue's host logic with a stand-in GPU step: states submitted from several threads are packed into
    chunks of at most 32 lanes, every job gets its own lanes' table back, tables are written as table.npz, a chunk
    whose solve fails fails its own jobs only, and a state of other profile shapes is refused at submit."""
    import threading
    from dataclasses import replace
    from types import SimpleNamespace

    from sosfuse2_amd import state_queue

    base = atmosphere.config1_state()
    scope = SimpleNamespace(state=base, context=lambda optics: SimpleNamespace(optics=optics))
    el, az = np.array([45.0, 15.0]), np.array([30.0, 90.0, 150.0])
    seen = []

    def fake_solve(scope_, items, el_, az_, ctx):
        seen.append(ctx.optics.W)
        if any(it.state.albedo == 0.99 for it in items):
            raise RuntimeError('E_NOT_CONVERGED: stand-in failure')
        out = []
        lane = 0
        for it in items:
            w = len(it.wavelengths)
            t = np.zeros((len(it.elevations), len(el_), len(az_), w, 3), 'f4')
            t[..., 0] = np.arange(lane, lane + w)[None, None, None, :] + 1
            out.append((t, dict(evaluations=7)))
            lane += w
        return out

    q = state_queue.StateQueue(scope, el, az, out_dir=tmp_path, linger=0.05, solve_chunk=fake_solve)
    jobs = []
    lock = threading.Lock()

    def client(k):
        for i in range(4):
            st = replace(base, name=f'c{k}_{i}', albedo=0.99 if (k, i) == (1, 2) else 0.1)
            j = q.submit(st, atmosphere.log_spaced_wavelengths(3 + 5 * i), [2.0, -3.0, -6.0][: 1 + i % 3])
            with lock:
                jobs.append(j)

    threads = [threading.Thread(target=client, args=(k,)) for k in range(3)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    with pytest.raises(ValueError, match='E_BATCH_PROFILE'):
        q.submit(replace(base, aerosol_height_km=2.0), [550.0], [-3.0])
    q.close()
    assert all(w <= 32 for w in seen) and len(seen) < len(jobs)
    failed = [j for j in jobs if j.error is not None]
    assert failed and all('E_NOT_CONVERGED' in str(j.error) for j in failed)
    assert any(j.item.state.albedo == 0.99 for j in failed)
    for j in jobs:
        if j.error is not None:
            with pytest.raises(RuntimeError):
                j.wait(1)
            continue
        table, info = j.wait(1)
        w = len(j.item.wavelengths)
        assert table.shape == (len(j.item.elevations), 2, 3, w, 3) and info['evaluations'] == 7
        lanes = table[0, 0, 0, :, 0]
        assert np.all(np.diff(lanes) == 1)  # its own contiguous lanes of the chunk
        axes, radiance = api.read_table(j.path)
        np.testing.assert_array_equal(radiance, table[..., 0])
    assert q.stats['states'] + q.stats['failed'] == len(jobs)
    with pytest.raises(RuntimeError, match='E_QUEUE_CLOSED'):
        q.submit(base, [550.0], [-3.0])


def test_stokes_of_pure_and_unpolarised_tensors():
    """api.stokes_from_tensor (device-side I, Q, U of the table) against closed forms: an isotropic tensor is
    unpolarised (I = 3, Q = U = 0); P = e e^T with e = e_v, e_h or (e_v + e_h)/sqrt 2 gives (1, 1, 0), (1, -1, 0) and
    (1, 0, 1); a view-parallel term d d^T adds to I only. Random tensors agree with the oracle's arrangement of the
    quadratic forms (oracle.solve.stokes)."""
    import torch

    from oracle.solve import stokes

    rng = np.random.default_rng(3)
    n = 64
    view = np.column_stack([rng.uniform(-1.5, 1.5, n), rng.uniform(-np.pi, np.pi, n)])
    el, az = view[:, 0], view[:, 1]
    eh = np.stack([-np.sin(az), np.cos(az), np.zeros(n)], -1)
    ev = np.stack([-np.sin(el) * np.cos(az), -np.sin(el) * np.sin(az), np.cos(el)], -1)
    d = np.cross(eh, ev)

    def comps(P):  # (xx, yy, zz, xz, xy, yz)
        return np.stack([P[:, 0, 0], P[:, 1, 1], P[:, 2, 2], P[:, 0, 2], P[:, 0, 1], P[:, 1, 2]], -1)[:, None, :]

    outer = lambda e: e[:, :, None] * e[:, None, :]
    cases = [(np.broadcast_to(np.eye(3), (n, 3, 3)), (3, 0, 0)), (outer(ev), (1, 1, 0)), (outer(eh), (1, -1, 0)),
             (outer((ev + eh) / np.sqrt(2)), (1, 0, 1)), (outer(ev) + 2 * outer(d), (3, 1, 0))]
    tv = torch.from_numpy(view)
    for P, want in cases:
        got = api.stokes_from_tensor(torch.from_numpy(comps(P)), tv).numpy()[:, 0]
        np.testing.assert_allclose(got, np.broadcast_to(want, got.shape), atol=1e-12)
    comp = rng.normal(size=(n, 5, 6))
    got = api.stokes_from_tensor(torch.from_numpy(comp), tv).numpy()
    np.testing.assert_allclose(got, stokes(comp, view), rtol=1e-12, atol=1e-12)


def test_seeded_solve_falls_back_to_the_cold_start():
    """Solver.solve with a start x0 that does not converge runs the iteration again from the cold start and counts
    both runs; without a start, or with fallback=False, the non-convergence is raised (E_NOT_CONVERGED)."""
    from types import SimpleNamespace

    from sosfuse2_amd.solver import Solver

    calls = []

    def iterate(max_eval, tolerance, x0):
        calls.append(x0)
        ok = x0 is None  # the seeded run stalls, the cold one converges
        return 'x', dict(evaluations=max_eval if not ok else 7, converged=[ok, True, ok]), ok

    sv = Solver.__new__(Solver)
    sv.s, sv.log = SimpleNamespace(W=3), None
    sv._iterate = iterate
    x, info = sv.solve(20, 1e-5, x0='start')
    assert calls == ['start', None] and x == 'x'
    assert info['fallback'] and info['evaluations_seeded'] == 20 and info['evaluations'] == 27
    calls.clear()
    with pytest.raises(RuntimeError, match='E_NOT_CONVERGED: 1/3'):
        sv.solve(20, 1e-5, x0='start', fallback=False)
    assert calls == ['start']
    sv._iterate = lambda m, t, x0: ('x', dict(evaluations=m, converged=[False, True, False]), False)
    with pytest.raises(RuntimeError, match='E_NOT_CONVERGED'):
        sv.solve(20, 1e-5)
