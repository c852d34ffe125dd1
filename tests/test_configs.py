"""GPU tests at the configurations' own workloads, through the HIP path.

- config0 exactly (60 levels, 550 nm, L = 16, Anderson depth 5, tolerance 1e-5, solar elevations -3 and -6, 4 x 6
  views): the solve and its I/Q/U table against the FP32 oracle's solve of the same inputs and the FP64 twin's table
  (tests/golden/config0_oracle.npz, written by tests/golden/make_config0.py).
- config1 at full size (150 x 75 cells, 30 wavelengths, L = 32, 25 solar elevations x 32 x 36 views): one warm pass
  (with a boundary) and first light on rays sampled across the grid against the FP32 oracle; the converged table's
  physical properties; table rays of two elevations recomputed by the FP64 oracle from the converged source.
- config2: the mixed-precision (split-FP16 SH) table of config1 against the FP32 table.
- config3: 64 seeded states of the batched distribution (AOD up to 0.5) on config1's grid, packed into chunks run as
  captured graphs: every state converges and matches the state solved alone.
"""

from pathlib import Path

import numpy as np
import pytest
import torch

from sosfuse2_amd import api, atmosphere, batch, geometry

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).with_name('golden') / 'config0_oracle.npz'


def rel(a, b):
    a, b = np.asarray(a, 'f8'), np.asarray(b, 'f8')
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def sub_pack(g, idx):
    n = len(g['d'])
    return {k: (v[idx] if isinstance(v, np.ndarray) and v.shape[:1] == (n,) else v) for k, v in g.items()}


# ------------------------------------------------------------------------------------------------------ config0
def test_config0_solve_and_table_against_the_oracle(config0_host):
    print('config0: solve and table', flush=True)
    ref = np.load(GOLDEN)
    scope = config0_host
    cfg = scope.cfg
    wl, hs, el, az = api.axes(cfg)
    np.testing.assert_array_equal(ref['solar_elevation'], hs)
    s = scope.context(atmosphere.lane_optics(atmosphere.config0_state(), wl))
    x, info = scope.solver(s, log=None).solve(60, float(cfg['anderson_tol']))
    assert int(cfg['anderson_depth']) == 5
    table, _ = api.make_table(s, x, hs, el, az, log=None)
    assert table.shape == (2, 4, 6, 1, 3)
    # same evaluation count as the oracle (the device rounds its composition factors once more: at most one apart)
    n_ref = int(ref['fp32_evaluations'])
    assert abs(info['evaluations'] - n_ref) <= 1
    src = x.cpu().numpy()[..., :1]
    assert rel(src, ref['fp32_source']) < (1e-5 if info['evaluations'] == n_ref else 3e-5)
    I, I32, I64 = table[..., 0], ref['fp32_table'][..., 0], ref['fp64_table'][..., 0]
    assert np.max(np.abs(I - I32) / I32) < 1e-4
    for k in (1, 2):
        assert np.max(np.abs(table[..., k] - ref['fp32_table'][..., k]) / I32) < 1e-4
    err64 = float(np.max(np.abs(I - I64) / I64))
    print(f'config0: {info["evaluations"]} evaluations (oracle {n_ref}); I within {err64:.2e} of the FP64 twin')
    assert err64 < 1e-4


# ------------------------------------------------------------------------------------------------------ config1
@pytest.fixture(scope='module')
def config1(config1_host):
    from types import SimpleNamespace

    print('config1: solve and table', flush=True)
    scope = config1_host
    wl, hs, el, az = api.axes(scope.cfg)
    s = scope.context(atmosphere.lane_optics(atmosphere.config1_state(), wl))
    sv = scope.solver(s, log=None)
    x, info = sv.solve(60, float(scope.cfg['anderson_tol']))
    table, _ = api.make_table(s, x, hs, el, az, log=None)
    return SimpleNamespace(scope=scope, s=s, solver=sv, x=x, info=info, table=table, wl=wl, hs=hs, el=el, az=az)


def _sample_rays(scope, per_cell=3, cells=16):
    """(bundle ray ids, host pack of those rays) for per_cell rays of cells spread over the grid."""
    packs = scope.packs
    ns = len(packs)
    off = np.r_[0, np.cumsum([len(p['d']) for p in packs])]
    picks, parts = [], []
    for cell in np.linspace(0, ns - 1, cells).astype(int):
        n = len(packs[cell]['d'])
        idx = np.linspace(0, n - 1, per_cell).astype(int)
        picks.append(off[cell] + idx)
        parts.append(sub_pack(packs[cell], idx))
    cat = geometry.concat_packs(parts)
    cat['end_columns'] = np.concatenate([p['end_columns'] for p in parts])
    return np.concatenate(picks), cat


def test_config1_warm_and_first_light_rays_against_the_oracle(config1):
    """One warm pass (the converged source as the field, a positive boundary) and first light over the whole config1
    bundle on the device; 48 rays of 16 cells spread over the grid recomputed by the FP32 oracle."""
    from oracle.transport import PAIRS4, transport

    c1 = config1
    s, k = c1.s, c1.s.kernel
    print('config1: sampled rays against the FP32 oracle', flush=True)
    db = c1.solver.geometry.bundle
    G = k.constants(s, db.nq).G
    rng = np.random.default_rng(5)
    bnd = (np.abs(rng.normal(size=(s.geo.nc, s.W))) * 1e-2).astype('f4')
    bd = torch.from_numpy(np.ascontiguousarray(np.concatenate([bnd, np.repeat(bnd[:, -1:], G - s.W, 1)], 1))).cuda()
    ids, cat = _sample_rays(c1.scope)
    field = c1.x.cpu().numpy()[..., : s.W]
    c = s.constants(s.geo.path_n)
    warm = k.radiance(db, s, c1.x, False, bd, 4).cpu().numpy()[ids, : s.W]
    first = k.radiance(db, s, None, True, None, 4).cpu().numpy()[ids, : s.W]
    assert rel(warm, transport(cat, c, field, False, bnd, PAIRS4)) < 2e-5
    assert rel(first, transport(cat, c, field, True, None, PAIRS4)) < 2e-5


def test_config1_table_is_physical(config1):
    t = config1.table
    assert t.shape == (25, 32, 36, 30, 3)
    I, Q, U = t[..., 0], t[..., 1], t[..., 2]
    assert np.isfinite(t).all()
    assert (I > 0).all()
    assert (np.hypot(Q, U) <= I * (1 + 1e-6)).all()
    assert all(config1.info['converged'])


def test_config1_table_rays_against_fp64(config1):
    """The table's accuracy against FP64: rays of two solar elevations recomputed by the FP64 oracle (first light and
    the multiple-scattering query of the converged source, Stokes I/Q/U, times the solar irradiance) agree with the
    table entries to 1e-4 of I."""
    from oracle.solve import stokes
    from oracle.transport import PAIRS6, transport

    c1 = config1
    s = c1.s
    print('config1: table rays against the FP64 oracle', flush=True)
    cq = s.constants(s.geo.query_n, np.float64)
    x64 = c1.x.cpu().numpy()[..., : s.W].astype('f8')
    ee, aa = np.meshgrid(c1.el, c1.az, indexing='ij')
    worst = 0.0
    for j in (6, 18):  # solar elevations 0 and -12 degrees
        h = float(c1.hs[j])
        cat = geometry.concat_packs(s.query_packs(h, ee, aa))
        idx = np.arange(0, len(cat['d']), 37)
        sub = sub_pack(cat, idx)
        tot = transport(sub, cq, x64, True, None, PAIRS6) + transport(sub, cq, x64, False, None, PAIRS6)
        v = stokes(tot, sub['view']) * s.solar[None, :, None]
        got = c1.table[j].reshape(-1, s.W, 3)[idx]
        err = float(np.max(np.abs(got - v) / np.abs(v[..., :1])))
        worst = max(worst, err)
    print(f'config1 table rays vs FP64 recomputation: {worst:.2e}')
    assert worst < 1e-4


def test_config2_mixed_table_against_fp32(config1):
    """config2: the same state with the SH contraction in split FP16 (FP32 accumulation). The query of the converged
    FP32 source in mixed precision agrees with the FP32 table to 1e-5 of I on every entry (the contraction's own
    error); the mixed solve, iterated to the same 1e-5 relative change, gives a table within 5e-5 of the FP32 one
    (both iterations stop within their tolerance of the fixed point)."""
    c1 = config1
    k = c1.s.kernel
    print('config2: mixed-precision solve and table', flush=True)
    I = c1.table[..., 0]
    k.mixed = True
    try:
        same, _ = api.make_table(c1.s, c1.x, c1.hs, c1.el, c1.az, log=None)
        s = c1.scope.context(atmosphere.lane_optics(atmosphere.config1_state(), c1.wl))
        x, info = c1.scope.solver(s, log=None).solve(60, 1e-5)
        table, _ = api.make_table(s, x, c1.hs, c1.el, c1.az, log=None)
    finally:
        k.mixed = False
    err_query = float(np.max(np.abs(same - c1.table) / I[..., None]))
    err = float(np.max(np.abs(table - c1.table) / I[..., None]))
    print(f'config2: query of the FP32 source within {err_query:.2e}; mixed solve ({info["evaluations"]} evaluations, '
          f'FP32 {c1.info["evaluations"]}) within {err:.2e} of the FP32 table')
    assert err_query < 1e-5
    assert err < 5e-5


# ------------------------------------------------------------------------------------------------------ config3
def test_config3_sixty_four_states(config1):
    """64 seeded states (AOD log-uniform 0.02-0.5, uneven wavelength and elevation counts) solved in packed chunks,
    each chunk's evaluation a captured graph: every state converges, and twelve of them (the thickest aerosol among
    them) match the state solved alone to 2e-5 of I."""
    c1 = config1
    scope = c1.scope
    items = batch.items_from_seed(64, 2026, top_km=120.0)
    assert max(it.state.aod550 for it in items) > 0.3
    print(f'config3: 64 states in {len(batch.pack(items, 32))} chunks', flush=True)
    got = batch.solve(scope, items, c1.el, c1.az, tolerance=1e-5, max_eval=60, lanes=32)
    assert all(all(info['converged']) for _, info in got)
    order = np.argsort([-it.state.aod550 for it in items])
    for i in list(order[:6]) + list(order[-6:]):
        it = items[i]
        s = scope.context(atmosphere.lane_optics(it.state, it.wavelengths))
        x, _ = scope.solver(s, log=None).solve(60, 1e-5)
        alone, _ = api.make_table(s, x, it.elevations, c1.el, c1.az, log=None)
        err = float(np.max(np.abs(got[i][0][..., 0] - alone[..., 0]) / alone[..., 0]))
        assert err < 2e-5, (i, it.state.aod550, err)
