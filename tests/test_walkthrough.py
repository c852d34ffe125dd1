"""The reference's frozen reduced table (`src/sosfuse2/tests/test_offline_cpu.py:24-57`): the walkthrough state
(clear_34 with O4 off, no stratospheric layer, the legacy exp8/triangle density, `src/sosfuse2/jobs/walkthrough.py:
183-196`) on the reduced grid (heights 0, 0.1, 8, 80 km, ten solar zenith nodes, nmu 1, nphi 2, path_n 4, L 16,
550 nm, tolerance 1e-3, sun table scale 4, query_n 6), solar elevations -3 and -6, 4 x 6 views.

tests/golden/walkthrough.json holds that state's coefficients (restated from the reference's JSON assets by
tests/golden/make_walkthrough.py) and the frozen values: table[:, ::3, ::5, 0] at rtol 1e-3 as the reference's test
asserts, and the sunward column of the reference's observed record (5 significant digits).

CPU: the FP64 oracle on the product's host geometry reproduces the frozen values (this pins the oracle's whole solve
and table end to end, through the reference's own golden vectors). GPU: the HIP path (device geometry, FP32
transport) reproduces them through api.create_configured."""

import json
from pathlib import Path

import numpy as np
import pytest

from sosfuse2_amd import api, atmosphere

DOC = json.loads((Path(__file__).with_name('golden') / 'walkthrough.json').read_text())


def walkthrough():
    """(cfg, state, optics, wavelengths, solar elevations, view elevations, relative azimuths) of the walkthrough."""
    st = DOC['settings']
    state = atmosphere.legacy_state(DOC['parameters'], DOC['no2_profile'], DOC['no2_profile_step_km'],
                                    DOC['depolarisation'], name='walkthrough')
    cfg = dict(h=st['h'], nr=len(st['h']), chi_grid='tiny', nchi=len(st['chi_deg']), nmu=st['nmu'], nphi=st['nphi'],
               path_n=st['path_n'], L=st['L'], top_km=st['top_km'], observer_km=st['observer_km'],
               query_n=st['query_n'], anderson_depth=st['anderson_depth'], anderson_tol=st['anderson_tol'],
               sun_scale=st['sun_scale'], wavelengths_nm=DOC['wavelength_nm'],
               solar_elevation_deg=st['solar_elevation_deg'], view_shape=st['view_shape'])
    wl, hs, el, az = api.axes(cfg)
    optics = atmosphere.optics_from_coefficients(state, wl, DOC['ext_per_km'], DOC['asca_per_km'], DOC['solar'])
    return cfg, state, optics, wl, hs, el, az


def check_frozen(I, evaluations):
    """I [2, 4, 6] against the frozen values; returns the largest relative deviation of the corners."""
    exp = DOC['expected']
    corners = np.asarray(exp['corners'])
    np.testing.assert_allclose(I[:, ::3, ::5], corners, rtol=exp['corners_rtol'], atol=0)
    np.testing.assert_allclose(I[:, :, 0], exp['sunward'], rtol=1e-4, atol=0)
    assert abs(I.min() / exp['minimum'] - 1) < 5e-4 and abs(I.max() / exp['maximum'] - 1) < 5e-4
    assert abs(evaluations - exp['evaluations']) <= 1
    return float(np.max(np.abs(I[:, ::3, ::5] / corners - 1)))


def test_walkthrough_settings_match_the_reference_grid():
    cfg, state, optics, wl, hs, el, az = walkthrough()
    grid = api.make_grid(cfg)
    np.testing.assert_allclose(np.rad2deg(grid.chi), DOC['settings']['chi_deg'], atol=1e-12)
    np.testing.assert_array_equal(el, [78.75, 56.25, 33.75, 11.25])
    np.testing.assert_array_equal(az, [15, 45, 75, 105, 135, 165])
    assert grid.ns == 40 and (16 + 1) * (16 + 2) // 2 == DOC['expected']['source_coefficients']
    assert optics.W == 1 and optics.ext[4, 0] == 0 and abs(optics.mix[0] - 0.97 / 1.015) < 1e-15
    with pytest.raises(ValueError, match='E_PROFILE_TABLE'):
        atmosphere.legacy_state(DOC['parameters'], DOC['no2_profile'][:100])


def test_oracle_reproduces_the_frozen_table():
    """FP64 oracle solve and table on the product's host geometry (sun table scale 4, 982 rays) against the
    reference's frozen values."""
    from oracle.solve import OracleSolver

    cfg, state, optics, wl, hs, el, az = walkthrough()
    scope = api.Scope(dict(cfg, workers=8), state, type('K', (), {'L': 16})())
    assert sum(len(g['d']) for g in scope.packs) == DOC['expected']['rays']
    ctx = scope.context(optics)
    orc = OracleSolver(ctx, np.float64, depth=int(cfg['anderson_depth']), packs=scope.packs)
    x, info = orc.solve(64, float(cfg['anderson_tol']))
    I = orc.table(x, hs, el, az, stokes_out=False)[..., 0]
    dev = check_frozen(I, info['evaluations'])
    last = info['history'][-1]
    print(f'walkthrough oracle: {info["evaluations"]} evaluations, corners within {dev:.2e}, last observer change '
          f'{max(last["observer_relative"]):.3g}')
    assert dev < 1e-4


@pytest.mark.gpu
def test_hip_path_reproduces_the_frozen_table(gpu):
    """api.create_configured (device geometry) + Solver.solve + make_table on the GPU against the frozen values."""
    cfg, state, optics, wl, hs, el, az = walkthrough()
    print('walkthrough: HIP solve and table', flush=True)
    s, scope = api.create_configured(None, cfg, wl, 'hip', state, sun_scale=int(cfg['sun_scale']), optics=optics)
    x, info = scope.solver(s, log=None).solve(64, float(cfg['anderson_tol']))
    table, _ = api.make_table(s, x, hs, el, az, log=None)
    assert table.shape == (2, 4, 6, 1, 3) and np.all(np.isfinite(table))
    dev = check_frozen(table[..., 0, 0], info['evaluations'])
    print(f'walkthrough HIP: {info["evaluations"]} evaluations, corners within {dev:.2e} of the frozen values')
    assert dev < 1e-4
