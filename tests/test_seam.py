"""The reference-shaped seam (sosfuse2_amd.seam.Kernel) driven by stand-ins that carry exactly the attribute names and
shapes the reference hands its kernel: a `ray_geometry.Bundle` (.cells, .nq, .rays, .arrays with the host CSR layout
of `src/sosfuse2/solve/ray_geometry.py:290-392`) and a solver context `s` with the attributes `Kernel.setup` reads
(`src/sosfuse2/solve/source_solver.py:1243-1339`). The atmosphere uses the reference's profile_at shapes
(`src/sosfuse2/backends/numpy_reference.py:187-223`): exp(-h / 8) air, exp(-h / H) aerosol, the triangle ozone, a
tabulated NO2 on the 0.5 km profile grid, O4 as air squared. The seam's results are compared with the oracle on the
same inputs: first light, a warm pass with a boundary, the ground-flux pass and the phase-2 query adapter.

CPU tests cover the seam's host pieces (the FP64 basis, the bundle conversion) and the oracle's profile kinds."""

import math
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from sosfuse2_amd import api, atmosphere, geometry, seam

STEP = 0.5
TOP = 80.0


def legacy_state():
    """clear-sky parameters with the reference's legacy profile shapes and a smooth tabulated NO2 (peak 1 at 25 km)."""
    z = np.arange(0.0, TOP + STEP / 2, STEP)
    no2 = np.exp(-0.5 * ((z - 25.0) / 6.0) ** 2) + 0.05 * np.exp(-z / 2.0)
    no2 /= no2.max()
    rows = np.zeros((6, len(z)))
    rows[3] = no2
    spec = atmosphere.ProfileSpec(
        kinds=(atmosphere.EXP, atmosphere.EXP, atmosphere.TRIANGLE, atmosphere.TABLE, atmosphere.AIR_SQUARED,
               atmosphere.ZERO),
        a=(8.0, 1.2, 15.0, STEP, 1.0, 1.0), b=(0.0, 0.0, 25.0, 0.0, 0.0, 0.0), table=tuple(tuple(r) for r in rows))
    st = atmosphere.State(name='legacy', top_km=TOP, aod550=0.15, aerosol_height_km=1.2, aerosol_g=0.7,
                          no2_column=5e15, albedo=0.2, profile_spec=spec)
    table = np.column_stack([np.exp(-z / 8.0), np.maximum(0, 1 - np.abs(z - 25) / 15), no2])
    return st, dict(table=table, mask=0, air_surface=1.0)


def reference_bundle(packs):
    """A stand-in of the reference's host Bundle (xp=numpy) built from cell packs."""
    cells = len(packs)
    edges = np.concatenate([g['edges'] for g in packs])
    lower = np.concatenate([g['lower'] for g in packs])
    ec, ev, fl = geometry.compact_edges(edges, lower)
    arrays = dict(
        cell_off=np.r_[0, np.cumsum([len(g['d']) for g in packs])].astype('i8'),
        starts=np.array([g['p'][0] for g in packs], 'f8'),
        targets=np.array([g['target'][0] for g in packs], 'f8'),
        dirs=np.concatenate([g['d'] for g in packs]),
        weights=np.concatenate([g['weight'] for g in packs]).astype('f4'),
        edge_off=np.r_[0, np.cumsum(ec)].astype('u4'),
        edges=ev.astype('f8'),
        lower=fl.astype('u2'),
        end_j=np.concatenate([g['end_index'] for g in packs]).astype('i4'),
        end_f=np.concatenate([g['end_fraction'] for g in packs]).astype('f4'),
        end_col=np.concatenate([g['end_columns'][:, :5] for g in packs]).astype('f4'),
        hits=np.concatenate([g['hits_ground'] for g in packs]).astype('u1'),
    )
    return SimpleNamespace(cells=cells, nq=packs[0]['nq'], rays=int(arrays['cell_off'][-1]), arrays=arrays)


def reference_context(ctx, density):
    """A stand-in of the reference's solver context `s` carrying the attributes Kernel.setup reads."""
    o = ctx.optics
    prof = atmosphere.profiles(ctx.state, ctx.geo.h)
    pools = ctx.pools
    return SimpleNamespace(
        W=o.W,
        geo=SimpleNamespace(h=ctx.geo.h, chi=ctx.geo.chi, ns=ctx.geo.ns,
                            sun=SimpleNamespace(h=ctx.sun.h, col=ctx.sun.col[..., :5])),
        ext=o.ext, asca=o.asca, hr=prof[:, 0], ha=prof[:, 1], delta=ctx.state.depolarisation,
        fuse_state={'parameters': {'albedo': float(ctx.state.albedo)}},
        goff=pools.off, gdir=pools.loc, gval=pools.values(o), gm=pools.tensor(o, ctx.geo.ns),
        density_optics=density,
    )


def _case(L=16, W=3):
    st, density = legacy_state()
    cfg = dict(nr=5, chi_grid='tiny', nmu=1, nphi=2, path_n=4, L=L, top_km=TOP, observer_km=0.1, query_n=6,
               anderson_depth=4, sun_scale=1, workers=1)
    scope = api.Scope(cfg, st, SimpleNamespace(L=L))
    wl = atmosphere.log_spaced_wavelengths(W, 420, 680)
    ctx = scope.context(atmosphere.lane_optics(st, wl))
    return st, density, scope, ctx


# ------------------------------------------------------------------------------------------------------ CPU
def test_seam_basis_matches_the_oracle_basis():
    from oracle.arith import basis

    mu = np.linspace(-0.98, 0.98, 9)
    for L in (8, 16, 32):
        np.testing.assert_allclose(seam.sh_basis(mu, L), basis(mu, L, np.float64), rtol=1e-10, atol=1e-12)


def test_host_bundle_conversion_keeps_the_reference_layout():
    st, density, scope, ctx = _case()
    ref = reference_bundle(scope.packs)
    host = seam._HostBundle(ref)
    ours = geometry.Bundle(scope.packs)
    for name in ('group_off', 'starts', 'dirs', 'weights', 'edge_off', 'edges', 'lower', 'end_j', 'end_f', 'hits',
                 'ray_group'):
        np.testing.assert_array_equal(getattr(host, name), getattr(ours, name), err_msg=name)
    np.testing.assert_allclose(host.dloc, ours.dloc, atol=1e-15)
    np.testing.assert_array_equal(host.end_col[:, :5], ours.end_col[:, :5])
    assert host.samples == ours.samples
    bad = SimpleNamespace(cells=ref.cells, nq=ref.nq, rays=ref.rays,
                          arrays=dict(ref.arrays, starts=np.zeros((ref.cells, 5), 'f4')))
    with pytest.raises(ValueError, match='E_SEAM_BUNDLE'):
        seam._HostBundle(bad)


def test_oracle_profile_kinds():
    """The oracle's triangle and table profiles against closed forms and numpy interpolation (the reference's
    interpolate32 is linear interpolation on the profile grid)."""
    from oracle import transport as ot

    st, density = legacy_state()
    spec = st.profiles
    h = np.linspace(0.0, TOP, 997)
    np.testing.assert_allclose(ot._profile(spec, 2, h, None, np.float64), np.maximum(0, 1 - np.abs(h - 25) / 15),
                               atol=1e-15)
    z = np.arange(0.0, TOP + STEP / 2, STEP)
    np.testing.assert_allclose(ot._profile(spec, 3, h, None, np.float64), np.interp(h, z, density['table'][:, 2]),
                               atol=1e-14)
    air = np.exp(-h / 8)
    np.testing.assert_allclose(ot._profile(spec, 4, h, air, np.float64), air * air, rtol=1e-15)
    assert np.all(ot._profile(spec, 5, h, air, np.float64) == 0)


# ------------------------------------------------------------------------------------------------------ GPU
@pytest.fixture(scope='module')
def seam_case(gpu):
    from oracle.solve import OracleSolver

    st, density, scope, ctx = _case()
    k = seam.Kernel(height=st.aerosol_height_km, g=st.aerosol_g, backend='hip', L=16,
                    domain=seam.Domain(top_km=TOP, step_km=STEP))
    s = reference_context(ctx, density)
    bundle = reference_bundle(scope.packs)
    ground = reference_bundle(scope.packs[: ctx.geo.nc])
    orc = OracleSolver(ctx, np.float32, depth=4, packs=scope.packs)
    return SimpleNamespace(st=st, scope=scope, ctx=ctx, k=k, s=s, bundle=bundle, ground=ground, orc=orc)


def rel(a, b):
    a, b = np.asarray(a, 'f8'), np.asarray(b, 'f8')
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.gpu
def test_seam_setup_constants(seam_case):
    c = seam_case.k.setup(seam_case.s, 4)
    cc = seam_case.ctx.constants(4)
    for key in ('u', 'wu', 'h', 'chi', 'hr', 'ha', 'ext', 'asca', 'gd', 'gv', 'gm'):
        np.testing.assert_allclose(np.asarray(c[key], 'f8'), np.asarray(cc[key], 'f8'), rtol=1e-6, atol=1e-30,
                                   err_msg=key)
    assert c['L'] == 16 and c['W'] == 3 and abs(c['mix'] - cc['mix'][0]) < 1e-7
    assert seam_case.k.setup(seam_case.s, 4) is c


@pytest.mark.gpu
def test_seam_launch_against_the_oracle(seam_case):
    from tests.test_gpu_parity import physical_field

    sc = seam_case
    k, s, ctx, orc = sc.k, sc.s, sc.ctx, sc.orc
    zero = np.zeros((ctx.geo.ns, ctx.rows, ctx.W), 'f4')
    m, f, q = k.launch(sc.bundle, s, None, True, 0)
    assert q is None and tuple(m.shape) == (ctx.geo.ns, 4 + k.K, ctx.W)
    mo, fo = orc.evaluate(zero, first=True, phase=0)
    assert rel(k.host(m), mo) < 1e-5
    assert rel(k.host(f), fo) < 1e-5
    field = physical_field(ctx, 4)
    bnd = np.abs(np.random.default_rng(2).normal(size=(ctx.geo.nc, ctx.W))).astype('f4') * 1e-2
    m, f, _ = k.launch(sc.bundle, s, torch.from_numpy(field).cuda(), False, 0, torch.from_numpy(bnd).cuda())
    mo, fo = orc.evaluate(field, phase=0, boundary=bnd.astype('f8'))
    assert rel(k.host(m), mo) < 1e-5
    m, f, _ = k.launch(sc.ground, s, field, False, 1)
    assert m is None
    _, fo = orc.evaluate(field, phase=1, ground_only=True)
    assert rel(k.host(f), fo) < 1e-5
    assert k.metrics['launches'] == 3


@pytest.mark.gpu
def test_seam_transport_adapter_against_the_oracle(seam_case):
    from oracle.transport import PAIRS4, transport
    from tests.test_gpu_parity import physical_field

    sc = seam_case
    k, s, ctx, orc = sc.k, sc.s, sc.ctx, sc.orc
    el, az, _ = ctx.box_directions()
    g = geometry.concat_packs(ctx.query_packs(-6.0, el, az))
    field = physical_field(ctx, 5)
    full = np.zeros(field.shape[:2] + (2 * ctx.W,), 'f4')
    full[..., ::2] = field
    for first in (True, False):
        out = k.transport(g, s, full, first)
        assert out.shape == (len(g['d']), ctx.W, 2, 4) and np.all(out[:, :, 1] == 0)
        ref = transport(g, orc.cq, field, first, None, PAIRS4)
        assert rel(out[:, :, 0], ref) < 2e-5, first


@pytest.mark.gpu
def test_seam_drives_the_reference_apply_loop_on_host_arrays(seam_case):
    """The reference Solver's NumPy branch (`src/sosfuse2/solve/source_iteration.py:123-216`: first light once, then
    per apply the ground pass, the boundary, the warm pass, moment + first-light moments, composition) driven through
    the seam with host arrays only (xp = numpy), two applies from a zero source, against the oracle's two applies."""
    from oracle.solve import compose

    sc = seam_case
    k, s, ctx, orc = sc.k, sc.s, sc.ctx, sc.orc
    assert k.xp is np
    nc, F = ctx.geo.nc, np.float32
    x = k.xp.zeros((ctx.geo.ns, ctx.rows, ctx.W), 'f4')
    m0, f0, _ = k.launch(sc.bundle, s, None, True, 0)
    assert isinstance(m0, np.ndarray) and isinstance(f0, np.ndarray)
    xo = x.copy()
    for _ in range(2):
        _, flux, _ = k.launch(sc.ground, s, x, False, 1)
        m, _, _ = k.launch(sc.bundle, s, x, False, 0, flux + f0[:nc])
        x = compose(m.astype('f8') + m0, orc.c, F)
        xo = orc.apply(xo)
    assert rel(x, xo) < 2e-5
