"""Pins of the CPU oracle: closed forms, invariants and the reference's own known-answer checks.

Each test names the reference check it restates or the mathematics that fixes the answer. A plausible mistake in the
oracle (a dropped term, a wrong sign or index, a transposed operand) fails at least one of them.
"""

import math

import numpy as np
import pytest

from oracle import arith
from oracle.arith import Arith, basis, cosine, flush_to_zero, kahan, tree_sum
from oracle.solve import anderson_weights, stokes
from oracle import transport as ot


def test_ftz_mode_flushes_subnormals():
    # numpy_reference.py:69-75: a normal input with a subnormal result, and a subnormal input, both give 0
    tiny = np.float32(np.finfo(np.float32).tiny)
    with flush_to_zero():
        assert tiny * np.float32(0.5) == 0
        assert (np.full(64, tiny, 'f4') * np.float32(0.5))[0] == 0
    assert tiny * np.float32(0.5) != 0  # restored on exit


def test_sh_low_orders():
    # checks/validate.py:198-201: Y00 = 1/sqrt(4 pi), Y10 at mu=0 is 0, Y10 at mu=-1 is -sqrt(3/(4 pi))
    y = basis(np.array([0.0, 1.0, -1.0]), 32)
    np.testing.assert_allclose(y[:, 0], 1 / math.sqrt(4 * math.pi), rtol=1e-7)
    assert abs(y[0, 1]) < 1e-7
    np.testing.assert_allclose(y[2, 1], -math.sqrt(3 / (4 * math.pi)), rtol=1e-7)


@pytest.mark.parametrize('L', [8, 16, 32])
def test_sh_orthonormal_on_the_sphere(L):
    # The cosine-only real harmonics Y_lm(mu) cos(m phi) (sqrt2 for m > 0) are orthonormal over the sphere: a wrong
    # recurrence constant, seed or sqrt2 factor breaks the Gram matrix.
    x, w = np.polynomial.legendre.leggauss(2 * L + 4)
    Y = basis(x, L).astype('f8')
    K = Y.shape[1]
    m = np.array([mm for l in range(L + 1) for mm in range(l + 1)])
    nphi = 4 * L + 8
    phi = (np.arange(nphi) + 0.5) * 2 * math.pi / nphi
    cm = np.cos(np.outer(phi, np.arange(L + 1)))  # [phi, m]
    az = cm[:, m].T @ cm[:, m] * (2 * math.pi / nphi)  # [K, K] azimuthal integrals
    gram = (Y * w[:, None]).T @ Y * az
    np.testing.assert_allclose(gram, np.eye(K), atol=2e-6)


def test_sh_basis_against_scipy_legendre():
    # Independent closed form: Y_lm = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!) P_lm(mu), times sqrt2 for m > 0
    from scipy.special import lpmv

    mu = np.linspace(-0.95, 0.95, 7)
    Y = basis(mu, 16).astype('f8')
    for l in range(17):
        for m in range(l + 1):
            k = l * (l + 1) // 2 + m
            norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - m) / math.factorial(l + m))
            ref = norm * lpmv(m, l, mu) * (math.sqrt(2) if m else 1)
            np.testing.assert_allclose(Y[:, k], ref, rtol=2e-5, atol=1e-7)


def test_chebyshev_cosines():
    cp = np.linspace(-1, 1, 11).astype('f4')
    cm = cosine(cp, 33).astype('f8')
    ref = np.cos(np.outer(np.arccos(cp.astype('f8')), np.arange(33)))
    np.testing.assert_allclose(cm, ref, atol=2e-5)


def test_tree_sum_of_33_lanes():
    # checks/validate.py:203-205: the 32-lane shuffle tree over 0..32 gives 528
    x = np.arange(33, dtype='f4')[None, :, None]
    assert tree_sum(x)[0, 0] == 528


def test_kahan_recovers_lost_bits():
    # Kahan's error term carries what a plain FP32 sum drops: 1 + 1e-8 * 1e4 exactly 1.0001 in compensated form
    total = np.zeros(1, 'f4')
    corr = np.zeros(1, 'f4')
    plain = np.float32(1)
    total[0] = 1
    for _ in range(10000):
        total, corr = kahan(total, corr, np.array([1e-8], 'f4'))
        plain = plain + np.float32(1e-8)
    assert plain == 1.0
    assert abs(float(total[0] - corr[0]) - 1.0001) < 2e-7


def test_double_single_coordinate_cancellation():
    # checks/validate.py:166-179: heights at Earth radius and on a near-tangent path to better than 2e-7 km
    a = Arith(np.float32)
    x = np.array([6371.0 + 0.000001, 6371.0 + 0.1, 6371.0 + 80.0])
    height = a.rounded(a.sub(a.split(x), (a.R, np.float32(0))))
    err = float(np.max(abs(height - (x - 6371.0))))
    p = np.array([[6371.1, 0.0, 0.0]])
    d = np.array([[-0.005, math.sqrt(1 - 0.005**2), 0.0]])
    t = -np.sum(p * d, axis=-1)
    q = a.add(a.split(p), a.mul(a.split(t[:, None]), a.split(d)))
    q2 = a.mul(q, q)
    rr = a.root(a.add(a.add((q2[0][:, 0], q2[1][:, 0]), (q2[0][:, 1], q2[1][:, 1])), (q2[0][:, 2], q2[1][:, 2])))
    h = a.rounded(a.sub(rr, (a.R, np.float32(0))))
    expected = np.linalg.norm(p + t[:, None] * d, axis=-1) - 6371.0
    err = max(err, float(np.max(abs(h - expected))))
    assert err < 2e-7


def test_log_blend_rule():
    # checks/validate.py:182-196: all corners valid -> geometric mean (3.0); two valid -> 2.5 sqrt(0.6); none -> linear
    F = np.float32
    traces = np.array([[1.0, 9.0, 1.0, 9.0], [1.0, 9.0, 0.0, 0.0], [0.0, 0.0, 0.0, 0.0]], 'f4')
    value = np.zeros((3, 1, 4), 'f4')
    value[:, 0, 0] = traces.mean(axis=1)
    logs = np.zeros((3, 1), 'f4')
    wg = np.zeros_like(logs)
    am = np.zeros_like(logs)
    for corner in range(4):
        ot.trace_accumulate(traces[:, corner, None], F(0.25), logs, wg, am, F)
    out = ot.trace_scale(value, logs, wg, am, F)
    np.testing.assert_allclose(out[:, 0, 0], [3.0, 2.5 * math.sqrt(0.6), 0.0], atol=2e-6)


def test_anderson_weights_equal_columns():
    # tests/test_transport_fp32.py:179-185: equal residual columns give weights 1/4
    f = np.full((32, 4), 3, 'f4')
    w = anderson_weights(f.T @ f)
    assert w.dtype == np.float32
    np.testing.assert_allclose(w, 0.25, rtol=1e-3)


def test_anderson_weights_minimise_the_residual():
    # The weights solve min |F w| subject to sum w = 1 (the Lagrange conditions of the regularised Gram system)
    rng = np.random.default_rng(3)
    F = rng.normal(size=(50, 3))
    w = anderson_weights(F.T @ F).astype('f8')
    G = F.T @ F
    ref = np.linalg.solve(G, np.ones(3))
    ref /= ref.sum()
    np.testing.assert_allclose(w, ref, rtol=1e-5)
    for _ in range(20):  # no other affine combination has a smaller residual
        v = ref + rng.normal(size=3) * 1e-3
        v /= v.sum()
        assert np.linalg.norm(F @ v) >= np.linalg.norm(F @ ref) - 1e-12


def test_anderson_guard_rejects_large_weights():
    # source_iteration.py:31: weights with max |w| >= 8 are rejected (the plain fixed-point step is taken).
    # G = [[1, a], [a, b]] gives w proportional to (b - a, 1 - a): a = 0.99, b = 0.9802 -> w = (-49, 50).
    assert anderson_weights(np.array([[1.0, 0.99], [0.99, 0.9802]])) is None
    w = anderson_weights(np.array([[1.0, 0.0], [0.0, 1.0]]))
    np.testing.assert_allclose(w, 0.5, rtol=1e-6)


def test_hg_moments_are_legendre_projections():
    # docs/MODEL.md:236-242: the HG phase function has Legendre moments (2l+1) g^l
    from sosfuse2_amd.atmosphere import hg_moments

    g = 0.6
    x, w = np.polynomial.legendre.leggauss(400)
    p = (1 - g * g) / (4 * math.pi * (1 + g * g - 2 * g * x) ** 1.5)
    chi = hg_moments(g, 12)
    for l in range(13):
        Pl = np.polynomial.legendre.Legendre.basis(l)(x)
        # p = sum chi_l P_l / (4 pi)  =>  chi_l = 2 pi (2l+1) int p P_l
        val = 2 * math.pi * (2 * l + 1) * np.sum(w * p * Pl)
        np.testing.assert_allclose(val, chi[l], rtol=1e-9)


def test_hg_phase_normalised():
    x, w = np.polynomial.legendre.leggauss(400)
    for g in (0.0, 0.6, 0.8):
        v = ot.hg(x[:, None].astype('f8'), np.array([g]), np.float64)
        assert abs(2 * math.pi * np.sum(w * v[:, 0]) - 1) < 1e-9


def test_stokes_of_rayleigh_single_scattering():
    # Rayleigh single scattering of unpolarised sunlight (docs/MODEL.md:214-221): with the depolarisation factor
    # D = (1 - delta)/(1 + delta/2) the degree of polarisation is D sin^2 T / (D (1 + cos^2 T) + 4 (1 - D)/3).
    # The direct-term tensor of the transport, D 3/(8pi) (P - e e)/2 + (1 - D) 1/(8pi) P, fed through the Stokes
    # projection, must give it for every view.
    D = (1 - 0.03) / (1 + 0.015)
    rng = np.random.default_rng(5)
    el = rng.uniform(0.05, 1.5, 40)
    az = rng.uniform(0, math.pi, 40)
    sun = np.array([math.cos(0.3), 0.0, math.sin(0.3)])  # sun direction in the observer frame
    d = np.stack([np.cos(el) * np.cos(az), np.cos(el) * np.sin(az), np.sin(el)], -1)
    P = np.eye(3)[None] - d[:, :, None] * d[:, None, :]
    e = sun[None] - d * (d @ sun)[:, None]
    T = D * 3 / (8 * math.pi) * (P - e[:, :, None] * e[:, None, :]) / 2 + (1 - D) / (8 * math.pi) * P
    comp = np.stack([T[:, 0, 0], T[:, 1, 1], T[:, 2, 2], T[:, 0, 2], T[:, 0, 1], T[:, 1, 2]], -1)[:, None, :]
    s = stokes(comp, np.stack([el, az], -1))[:, 0]
    cos = d @ sun
    dop = np.hypot(s[:, 1], s[:, 2]) / s[:, 0]
    ref = D * (1 - cos**2) / (D * (1 + cos**2) + 4 * (1 - D) / 3)
    np.testing.assert_allclose(dop, ref, rtol=1e-10)
    # the phase function: I = 3/(16 pi) D (1 + cos^2) + (1 - D)/(4 pi), normalised over the sphere
    np.testing.assert_allclose(s[:, 0], D * 3 / (16 * math.pi) * (1 + cos**2) + (1 - D) / (4 * math.pi), rtol=1e-12)


def _cpu_context(L=8, nr=5, nmu=1, nphi=2, path_n=4, state=None, wavelengths=(450.0, 550.0, 650.0)):
    """A small context without a GPU: the host geometry with a stand-in kernel that only carries L."""
    from types import SimpleNamespace

    from sosfuse2_amd import api, atmosphere

    cfg = dict(nr=nr, chi_grid='tiny', nmu=nmu, nphi=nphi, path_n=path_n, L=L, top_km=100.0, observer_km=0.1,
               query_n=6, anderson_depth=4, sun_scale=1, workers=1)
    st = state or atmosphere.config1_state()
    scope = api.Scope(cfg, st, SimpleNamespace(L=L))
    return scope.context(atmosphere.lane_optics(st, np.asarray(wavelengths))), scope


def _truth_tau(g, c, e0, e1):
    """Independent FP64 optical depth along each ray between path coordinates e0 and e1: Gauss-Legendre 48 on ten
    equal pieces, heights from |p + t d| - R in FP64, the same profile model."""
    x, w = np.polynomial.legendre.leggauss(48)
    out = np.zeros((len(e0), c['W']))
    for piece in range(10):
        a = e0 + (e1 - e0) * piece / 10
        b = e0 + (e1 - e0) * (piece + 1) / 10
        for xi, wi in zip(x, w):
            t = (a + b) / 2 + (b - a) / 2 * xi
            q = g['p'] + t[:, None] * g['d']
            h = np.clip(np.linalg.norm(q, axis=1) - 6371.0, 0.0, c['top'])
            prof, st = ot._profiles(h, c, np.float64)
            out += ((b - a) / 2 * wi)[:, None] * ot._extinction(prof, st, c, np.float64)
    return out


def _prefix_errors(path_n):
    ctx, scope = _cpu_context(path_n=path_n)
    c = ctx.constants(path_n, np.float64)
    g = geometry_concat(scope.packs[:6])
    ops = ot._Ops(np.float64)
    pre = ot.Simpson(g, c, ops)
    st = np.sqrt(np.maximum(0.0, 1 - g['d'][:, 2] ** 2))
    closure = np.zeros((len(g['d']), c['W']))
    for seg in range(g['edges'].shape[1] - 1):
        pre.begin(seg)
        for step in range(len(c['u'])):
            q = ot.sample(g, seg, step, c, ops, st, ot.PAIRS4)
            if not q['valid'].any():
                continue
            mid = pre.node(q)
            beta = ot._extinction(q['prof'], q['st'], c, np.float64)
            closure += np.where(q['valid'][:, None], q['ds'][:, None] * beta * np.exp(-mid), 0.0)
        pre.finish()
    e = g['edges']
    valid = (e[:, 1:] - e[:, :-1]) > 1e-8
    truth = np.zeros_like(pre.tau)
    for seg in range(e.shape[1] - 1):
        truth += np.where(valid[:, seg, None], _truth_tau(g, c, e[:, seg], e[:, seg + 1]), 0.0)
    tau_err = float(np.max(np.abs(pre.tau - truth) / np.maximum(truth, 1e-12)))
    closure_err = float(np.max(np.abs(closure - (1.0 - np.exp(-pre.tau)))))
    return tau_err, closure_err


def test_simpson_prefix_against_fp64_integral_and_closure():
    """The Simpson optical-depth prefix driven as the transport drives it (begin, node(sample), finish) against an
    independent FP64 integral of the same profiles (the truth of the reference's test_simpson_prefix.py), and the
    closure sum_k ds beta e^-tau_k = 1 - e^-tau_end of the attenuation weights (its E4 check). Composite Simpson
    converges at fourth order: each doubling of the nodes per segment divides the error by about 16 (asserted > 10),
    and 32 nodes reach the FP64 truth to 5e-6. A lost factor 4 on the midpoint or a wrong panel weight stops the
    convergence at O(1) error; a wrong node position or ds breaks the closure."""
    errs = [_prefix_errors(n) for n in (8, 16, 32)]
    for (a, _), (b, _) in zip(errs, errs[1:]):
        assert a / b > 10.0
    assert errs[-1][0] < 5e-6
    assert errs[-1][1] < 1e-5


def geometry_concat(packs):
    from sosfuse2_amd import geometry

    return geometry.concat_packs(packs)


def test_anderson_iteration_known_affine_fixed_point():
    """The reference's D1 check (src/sosfuse2/tests/test_iteration.py:25-61) on the oracle's iteration: a diagonal
    affine map forcing + q x with a known fixed point forcing / (1 - q), one wavelength settling at the first apply
    (q = 0) and staying frozen, and the non-convergence signal."""
    from types import SimpleNamespace

    from oracle.solve import OracleSolver

    waves, ns, rows = 15, 2, 571
    q = np.linspace(0.25, 0.75, waves, dtype='f4')[None, None, :]
    q[:, :, 0] = 0
    forcing = np.linspace(0.01, 0.2, ns * rows * waves, dtype='f4').reshape(ns, rows, waves)
    expected = forcing / (1 - q)

    def solver(apply):
        o = OracleSolver.__new__(OracleSolver)
        o.ctx = SimpleNamespace(geo=SimpleNamespace(ns=ns), rows=rows)
        o.c = dict(W=waves)
        o.F = np.float32
        o.depth = 4
        o.apply = apply
        o.boxes = lambda field: np.broadcast_to(field[0, 0], (5, 7, waves)).copy()
        return o

    field, stats = solver(lambda x: forcing + q * x).solve(16, 0.005)
    relative = float(np.max(np.abs(field - expected) / expected))
    assert relative <= float(q.max() / (1 - q.max()) * 0.005) + 1e-6
    first_done = next(i for i, row in enumerate(stats['history']) if row['converged'][0])
    assert first_done < len(stats['history']) - 1
    assert all(row['source_relative'][0] == 0 and row['observer_relative'][0] == 0
               for row in stats['history'][first_done + 1:])
    with pytest.raises(RuntimeError, match='E_NOT_CONVERGED'):
        solver(lambda x: x + forcing).solve(2, 0.005)


@pytest.mark.parametrize('nmu,nphi,tol', [(4, 8, 1e-5), (8, 16, 1e-10)])
def test_projection_of_isotropic_light(nmu, nphi, tol):
    """Projection of a unit isotropic radiance (trace 1, tensor diag 1/3) through the incident quadrature: the solid
    angle converges to 4 pi (the reference records the same sum as solid_angle_max_relative,
    src/sosfuse2/solve/engine_core.py:247, for its identical sinh-graded quadrature), the (0, 0) moment to
    4 pi Y00 = sqrt(4 pi), each diagonal tensor moment to 4 pi / 3, and the ground flux to
    albedo / pi * integral of max(0, cos) = albedo. A wrong weight, a Y00 or trace slip, or the flux taken on the
    wrong axis misses these by O(1)."""
    ctx, scope = _cpu_context(nmu=nmu, nphi=nphi)
    c = ctx.constants(ctx.geo.path_n, np.float64)
    for cell in (0, 7, len(scope.packs) // 2, len(scope.packs) - 1):
        g = scope.packs[cell]
        rad = np.zeros((len(g['d']), c['W'], 4))
        rad[:, :, :3] = 1.0 / 3.0
        out, ground = ot.project(g, rad, c)
        assert abs(g['weight'].sum() / (4 * math.pi) - 1) < tol
        assert np.allclose(out[4], math.sqrt(4 * math.pi), rtol=tol, atol=0)
        for k in range(3):
            assert np.allclose(out[k], 4 * math.pi / 3, rtol=tol, atol=0)
        assert np.allclose(out[3], 0.0, atol=1e-12)
        assert np.allclose(ground, c['albedo'], rtol=max(tol, 1e-3), atol=0)


def test_mlp_oracle_gradients_by_finite_differences():
    """The network oracle's backward pass (FP64, no rounding) against central differences of its loss, for every
    weight and bias of a small 4-layer GELU network; and gelu' against a difference quotient. A transposed operand,
    a lost GELU derivative or a wrong 2/n factor fails by O(1)."""
    from oracle import mlp

    rng = np.random.default_rng(5)
    sizes = [7, 6, 5, 6, 4]
    params = [(rng.normal(size=(o, i)) / math.sqrt(i), rng.normal(size=o) * 0.1) for i, o in zip(sizes, sizes[1:])]
    x = rng.normal(size=(9, sizes[0]))
    t = rng.normal(size=(9, sizes[-1]))
    wt = rng.uniform(0.5, 2.0, size=sizes[-1])
    y, cache = mlp.forward(params, x)
    grads = mlp.backward(params, cache, y, t, wt)
    h = 1e-6
    for li, (W, b) in enumerate(params):
        for arr, g in ((W, grads[li][0]), (b, grads[li][1])):
            num = np.zeros_like(arr)
            for idx in np.ndindex(arr.shape):
                old = arr[idx]
                arr[idx] = old + h
                lp = mlp.loss(mlp.forward(params, x)[0], t, wt)
                arr[idx] = old - h
                lm = mlp.loss(mlp.forward(params, x)[0], t, wt)
                arr[idx] = old
                num[idx] = (lp - lm) / (2 * h)
            assert np.allclose(g, num, rtol=1e-6, atol=1e-9)
    z = np.linspace(-4, 4, 81)
    assert np.allclose(mlp.gelu_grad(z), (mlp.gelu(z + 1e-6) - mlp.gelu(z - 1e-6)) / 2e-6, rtol=1e-7, atol=1e-9)
    # gelu closed values: gelu(0) = 0, gelu(z) -> z for large z, gelu(-z) = gelu(z) - z
    assert mlp.gelu(0.0) == 0.0 and abs(mlp.gelu(10.0) - 10.0) < 1e-12
    assert np.allclose(mlp.gelu(-z), mlp.gelu(z) - z, atol=1e-15)


def test_bf16_rounding_ties_to_even():
    """bf16 keeps the top 16 bits with round-to-nearest-even: 1 + 2^-8 (a tie) rounds to 1, 1 + 3 2^-8 to 1 + 2^-6,
    and the result always has zero low bits."""
    from oracle import mlp

    x = np.array([1.0, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, -2.5, 3.0e38, 1e-30], np.float32)
    r = mlp.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1 + 2.0 ** -6 and r[3] == -2.5
    assert np.all((r.view(np.uint32) & 0xFFFF) == 0)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8)


def test_adam_first_step_moves_by_the_learning_rate():
    """Adam's bias-corrected first step is lr * g / (|g| + eps): every parameter moves by about lr against its
    gradient's sign, whatever the gradient's scale."""
    from oracle import mlp

    p = np.array([1.0, -2.0, 0.5])
    g = np.array([1e-3, -5.0, 2e2])
    m, v = np.zeros(3), np.zeros(3)
    mlp.adam(p, g, m, v, 1, lr=0.01)
    assert np.allclose(p, [1.0 - 0.01, -2.0 + 0.01, 0.5 - 0.01], rtol=0, atol=1e-7)


def _sun_truth(r, mu, c, top):
    """FP64 column integrals of the six unit profiles from a point at radius r along a direction of cosine mu to the
    z axis (the sun), up to the top of the atmosphere (scipy quad, split at the closest approach)."""
    from scipy.integrate import quad

    R = 6371.0
    s_top = -r * mu + math.sqrt(r * r * mu * mu - r * r + (R + top) ** 2)
    st = math.sqrt(max(0.0, 1 - mu * mu))
    split = [s for s in (-r * mu,) if 0 < s < s_top]

    def height(s):
        return math.hypot(r * st * 0 + s * st, r + s * mu) - R

    def prof(s, k):
        h = min(max(height(s), 0.0), top)
        p, sp = ot._profiles(np.array([h]), c, np.float64)
        return float(p[0, k]) if k < 5 else float(sp[0])

    return np.array([quad(prof, 0.0, s_top, args=(k,), points=split or None, limit=400, epsabs=0, epsrel=1e-10)[0]
                     for k in range(6)])


def test_sun_columns_against_fp64_path_integrals():
    """The oracle's sun columns (table lookup of sun_columns, numpy_reference.py:431-458, on the host-built sun
    table) against FP64 path integrals of the same profiles toward the sun: at table nodes the only error is the
    table's quadrature (< 1e-6 relative); between nodes the bilinear interpolation of the 81 x 401 table (up to 13 %
    on the small columns of grazing paths near the top, so the check there is 25 %). A swapped component, a wrong
    solar coordinate x or lit flag fails by O(1)."""
    ctx, scope = _cpu_context()
    c = ctx.constants(ctx.geo.path_n, np.float64)
    top = c['top']
    R = 6371.0
    ops = ot._Ops(np.float64)
    sh, scol = c['sun_h'], c['sun_col']
    nx = scol.shape[1]
    rng = np.random.default_rng(4)
    # table nodes
    for i in (0, 5, len(sh) // 3, len(sh) - 3):
        for j in (1, nx // 7, nx // 2, nx - 1):
            r = R + float(sh[i])
            mg = -math.sqrt(max(0.0, 1 - (R / r) ** 2))
            x = j / (nx - 1)
            mu = mg + (1 - mg) * x * x
            q = dict(h=np.array([sh[i]]), r=(np.array([r]), np.zeros(1)), qz=(np.array([mu * r]), np.zeros(1)))
            col, lit = ot.sun_columns(q, c, ops)
            assert lit[0]
            truth = _sun_truth(r, mu, c, top)
            big = np.abs(truth) > 1e-9 * np.abs(truth).max()
            assert np.allclose(col[0][big], truth[big], rtol=1e-6, atol=0), (i, j)
    # between nodes
    for _ in range(12):
        h = float(rng.uniform(0.5, top - 1))
        r = R + h
        mg = -math.sqrt(max(0.0, 1 - (R / r) ** 2))
        mu = float(mg + (1 - mg) * rng.uniform(0.05, 1.0) ** 2)
        q = dict(h=np.array([h]), r=(np.array([r]), np.zeros(1)), qz=(np.array([mu * r]), np.zeros(1)))
        col, lit = ot.sun_columns(q, c, ops)
        truth = _sun_truth(r, mu, c, top)
        big = np.abs(truth) > 1e-3 * np.abs(truth).max()
        assert lit[0]
        assert np.allclose(col[0][big], truth[big], rtol=0.25, atol=0), (h, mu)
    # below the horizon: not lit
    r = R + 10.0
    mg = -math.sqrt(1 - (R / r) ** 2)
    q = dict(h=np.array([10.0]), r=(np.array([r]), np.zeros(1)), qz=(np.array([(mg - 0.01) * r]), np.zeros(1)))
    assert not ot.sun_columns(q, c, ops)[1][0]


def test_first_light_rayleigh_single_scattering():
    """The first-light transport assembled by the oracle (node sampling, Simpson attenuation, sun columns, the
    Rayleigh tensor with depolarisation, the corner blend) against the textbook single-scattering integral of a
    pure Rayleigh atmosphere, evaluated independently in FP64:
        I = int_0^L beta_R(s) exp(-tau_view(0, s)) exp(-tau_sun(s)) [mix 3 (1 + cos^2) / (16 pi) + (1 - mix) / (4 pi)] ds
    with cos = d_z (the sun on the z axis), for rays of cells in full sun. The trace xx + yy + zz of the oracle's
    tensor radiance must agree to 5e-3 (its node quadrature and the sun table's interpolation); a lost factor,
    a wrong depolarisation weight or a missing attenuation fails by O(1)."""
    from scipy.integrate import quad

    from sosfuse2_amd import atmosphere

    st = atmosphere.State(aod550=0.0, ozone_du=0.0, no2_column=0.0, o4=False, sulfate_aod550=0.0, albedo=0.0)
    ctx, scope = _cpu_context(state=st, path_n=8, nmu=2, nphi=4, wavelengths=(550.0,))
    c = ctx.constants(8, np.float64)
    extR = float(c['ext'][0][0])
    mix = float(c['mix'][0])
    top, R, H = c['top'], 6371.0, st.air_height_km
    air = lambda x: math.exp(-min(max(np.linalg.norm(x) - R, 0.0), top) / H)
    nc = ctx.geo.nc
    checked = 0
    for ih in (1, 3):
        for ic in (0, 1):  # chi = 0 and 60 degrees: every node sunlit
            g = scope.packs[ih * nc + ic]
            pick = np.linspace(0, len(g['d']) - 1, 5).astype(int)
            sub = {k: (v[pick] if isinstance(v, np.ndarray) and v.shape[:1] == (len(g['d']),) else v)
                   for k, v in g.items()}
            tens = ot.transport(sub, c, np.zeros((ctx.geo.ns, ctx.rows, 1)), True)
            trace = tens[:, 0, 0] + tens[:, 0, 1] + tens[:, 0, 2]
            for r in range(len(pick)):
                p, d = sub['p'][r], sub['d'][r]
                L = float(sub['edges'][r][-1])
                if L <= 0:
                    continue

                def sun_col(x):
                    rr = np.linalg.norm(x)
                    mu = x[2] / rr
                    s_top = -rr * mu + math.sqrt(rr * rr * mu * mu - rr * rr + (R + top) ** 2)
                    return quad(lambda t: air(x + t * np.array([0, 0, 1.0])), 0, s_top, limit=200,
                                epsrel=1e-10)[0]

                phase = mix * 3 * (1 + d[2] ** 2) / (16 * math.pi) + (1 - mix) / (4 * math.pi)
                f = lambda s: extR * air(p + s * d) * math.exp(
                    -extR * quad(lambda u: air(p + u * d), 0, s, limit=200, epsrel=1e-10)[0]
                    - extR * sun_col(p + s * d)) * phase
                truth = quad(f, 0, L, limit=200, epsrel=1e-8)[0]
                assert abs(trace[r] - truth) <= 5e-3 * truth, (ih, ic, r, trace[r], truth)
                checked += 1
    assert checked >= 12
