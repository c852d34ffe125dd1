"""Synthetic atmosphere states: height profiles, extinction coefficients and phase-function moments.

The transport reads an atmosphere only through the quantities the reference's kernel constants carry
(`src/sosfuse2/solve/source_solver.py:1288-1330`): per-component extinction coefficients ext[5, W] at unit profile
value, the aerosol scattering coefficient asca[W], the component profiles at a height, the Henyey-Greenstein asymmetry,
the ground albedo, and for a stratospheric layer its extinction, scattering, Legendre moments and profile.
The component order is the reference's (`src/sosfuse2/atmosphere/components.py:23`): air, tropospheric aerosol,
ozone, NO2, O4.

Profiles here are the analytic shapes of the synthetic states (exponential air and aerosol, Gaussian ozone, NO2 and
sulfate, O4 as the square of the normalised air density), defined on the whole calculation domain. The reference's
profile_at (`src/sosfuse2/backends/numpy_reference.py:187-223`) evaluates its own shapes (exp(-h/8), exp(-h/H), a
triangle ozone, a tabulated NO2); the extinction model `ext[k] * profile_k(h)` is the same
(`numpy_reference.py:226-233`).

Every lane of a solve carries its own optical state, so one launch can hold wavelengths of several states (batching).
"""

from dataclasses import dataclass, field, replace
import math

import numpy as np

NCOMP = 5
COMPONENTS = ('air', 'tropo_aerosol', 'o3', 'no2', 'o4')
LOSCHMIDT_DU = 2.687e16  # molecules cm^-2 per Dobson unit
KM_CM = 1e5


# Profile kinds (the device's, include/sosfuse2_hip.h): exponential (a = scale height), Gaussian (a = sigma,
# b = centre), triangle (a = half width, b = centre), table (a = step, values from height 0), the square of the air
# profile over its surface value (a = surface value; O4), zero.
EXP, GAUSS, TRIANGLE, TABLE, AIR_SQUARED, ZERO = 1, 2, 3, 4, 5, 6


@dataclass(frozen=True)
class ProfileSpec:
    """Shapes of the five component profiles (air, aerosol, O3, NO2, O4) and the sulfate layer: kinds[k] with
    parameters a[k], b[k]; `table` holds one row per component (tuples, so the spec is hashable) on a uniform grid of
    step a[k] from height 0 for the TABLE kinds."""

    kinds: tuple
    a: tuple
    b: tuple
    table: tuple = None

    def table_array(self):
        return None if self.table is None else np.asarray(self.table, 'f8')


def evaluate_profile(spec, k, h, air=None, dtype=np.float64):
    """Profile k of a spec at heights h (>= 0) in dtype; air: the air profile for AIR_SQUARED."""
    F = np.dtype(dtype).type
    h = np.asarray(h, dtype)
    kind, a, b = spec.kinds[k], F(spec.a[k]), F(spec.b[k])
    if kind == EXP:
        return np.exp(-h / a)
    if kind == GAUSS:
        t = (h - b) / a
        return np.exp(F(-0.5) * (t * t))
    if kind == TRIANGLE:
        return np.maximum(F(0), F(1) - np.abs(h - b) / a)
    if kind == TABLE:
        row = np.asarray(spec.table[k], dtype)
        x = np.maximum(h, F(0)) * (F(1) / a)
        i = np.minimum(x.astype('i8'), len(row) - 2)
        f = x - i.astype(dtype)
        return row[i] + f * (row[i + 1] - row[i])
    if kind == AIR_SQUARED:
        q = air if air is not None else evaluate_profile(spec, 0, h, dtype=dtype)
        return (q * q) / (a * a)
    if kind == ZERO:
        return np.zeros_like(h)
    raise ValueError(f'E_PROFILE_KIND: {kind}')


@dataclass(frozen=True)
class State:
    """One atmosphere parameter set (a synthetic state of the configurations)."""

    name: str = 'config0'
    top_km: float = 100.0
    rayleigh_tau550: float = 0.097
    air_height_km: float = 8.0
    depolarisation: float = 0.03  # delta; mix = (1 - delta) / (1 + delta / 2) (source_solver.py:1294)
    aod550: float = 0.1
    aerosol_height_km: float = 1.2
    aerosol_g: float = 0.7
    aerosol_ssa: float = 0.95
    angstrom: float = 1.3
    ozone_du: float = 300.0
    ozone_center_km: float = 22.0
    ozone_sigma_km: float = 5.0
    no2_column: float = 0.0  # molecules cm^-2
    no2_center_km: float = 25.0
    no2_sigma_km: float = 5.0
    o4: bool = False
    sulfate_aod550: float = 0.0
    sulfate_center_km: float = 20.0
    sulfate_sigma_km: float = 3.0
    sulfate_g: float = 0.6
    sulfate_ssa: float = 1.0
    sulfate_angstrom: float = 1.0
    albedo: float = 0.1
    profile_spec: ProfileSpec = None  # None: the analytic shapes of the fields above

    @property
    def strato(self):
        return self.sulfate_aod550 > 0

    @property
    def profiles(self):
        """The ProfileSpec of this state: its own, or exponential air and aerosol, Gaussian ozone, NO2 and sulfate,
        O4 as the square of the air profile."""
        if self.profile_spec is not None:
            return self.profile_spec
        return ProfileSpec(kinds=(EXP, EXP, GAUSS, GAUSS, AIR_SQUARED, GAUSS),
                           a=(self.air_height_km, self.aerosol_height_km, self.ozone_sigma_km, self.no2_sigma_km, 1.0,
                              self.sulfate_sigma_km),
                           b=(0.0, 0.0, self.ozone_center_km, self.no2_center_km, 0.0, self.sulfate_center_km))

    @property
    def mix(self):
        d = self.depolarisation
        return (1 - d) / (1 + d / 2)


def config0_state():
    """Small configuration: Rayleigh, tropospheric aerosol, ozone, ground albedo 0.1, top 100 km."""
    return State(name='config0')


def config1_state():
    """Full configuration: config0's aerosol plus a sulfate layer, NO2 and O4, top 120 km."""
    return State(name='config1', top_km=120.0, no2_column=5e15, o4=True, sulfate_aod550=0.02)


def sample_states(n, seed=0, top_km=120.0, max_wavelengths=30, max_elevations=25):
    """The batched distribution: AOD log-uniform 0.02-0.5, g U(0.6, 0.8), sulfate AOD U(0, 0.05), ozone U(250, 400) DU,
    albedo U(0, 0.3); per state an uneven wavelength count U{1..30} and solar-elevation count U{2..25}."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        s = State(
            name=f'batch{seed}_{i}',
            top_km=top_km,
            aod550=float(np.exp(rng.uniform(np.log(0.02), np.log(0.5)))),
            aerosol_g=float(rng.uniform(0.6, 0.8)),
            sulfate_aod550=float(rng.uniform(0.0, 0.05)),
            ozone_du=float(rng.uniform(250, 400)),
            albedo=float(rng.uniform(0.0, 0.3)),
            no2_column=5e15,
            o4=True,
        )
        nw = int(rng.integers(1, max_wavelengths + 1))
        ne = int(rng.integers(2, max_elevations + 1))
        out.append((s, nw, ne))
    return out


def gaussian(h, center, sigma):
    t = (np.asarray(h, 'f8') - center) / sigma
    return np.exp(-0.5 * t * t)


def profiles(state, h):
    """Component profiles [n, 5] at heights h (km), FP64, from the state's ProfileSpec. With the analytic shapes air
    and aerosol are 1 at the ground, ozone and NO2 peak at 1 and O4 is the square of the air profile
    (`numpy_reference.py:219-220`)."""
    h = np.maximum(np.asarray(h, 'f8'), 0.0)
    spec = state.profiles
    air = evaluate_profile(spec, 0, h)
    return np.stack([air] + [evaluate_profile(spec, k, h, air) for k in (1, 2, 3, 4)], -1)


def strato_profile(state, h):
    """Sulfate layer profile, peak 1 (`numpy_reference.py:423-428` with norm 1)."""
    return evaluate_profile(state.profiles, 5, np.maximum(np.asarray(h, 'f8'), 0.0))


def _column_exp(H, top):
    return H * (-math.expm1(-top / H))


def _column_gauss(c, s, top):
    a = (0 - c) / (s * math.sqrt(2))
    b = (top - c) / (s * math.sqrt(2))
    return s * math.sqrt(math.pi / 2) * (math.erf(b) - math.erf(a))


def ozone_cross_section(wl):
    """Synthetic smooth O3 cross-section (cm^2): a Chappuis band near 602 nm plus the Huggins tail."""
    wl = np.asarray(wl, 'f8')
    return 4.7e-21 * np.exp(-0.5 * ((wl - 602.0) / 55.0) ** 2) + 1e-19 * np.exp(-(wl - 330.0) / 12.0) + 2e-23


def no2_cross_section(wl):
    """Synthetic smooth NO2 cross-section (cm^2), broad maximum near 420 nm."""
    wl = np.asarray(wl, 'f8')
    return 6e-19 * np.exp(-0.5 * ((wl - 420.0) / 60.0) ** 2)


def o4_cross_section(wl):
    """Synthetic O4 collision-induced absorption (cm^5 molecule^-2): Gaussian bands at 477, 532, 577 and 630 nm."""
    wl = np.asarray(wl, 'f8')
    bands = ((477.0, 6.3e-46, 4.0), (532.0, 1.0e-46, 4.0), (577.0, 1.1e-45, 5.0), (630.0, 7.5e-46, 5.0))
    return sum(a * np.exp(-0.5 * ((wl - c) / w) ** 2) for c, a, w in bands)


def solar_irradiance(wl):
    """Synthetic top-of-atmosphere irradiance (W m^-2 nm^-1): a 5778 K Planck curve scaled to 1.85 at 550 nm."""
    wl = np.asarray(wl, 'f8') * 1e-9

    def planck(x):
        return 1.0 / (x**5 * np.expm1(0.014387769 / (x * 5778.0)))

    return 1.85 * planck(wl) / planck(550e-9)


@dataclass
class Optics:
    """Per-lane optical constants of one solve chunk (FP64 host values; the device receives FP32 copies).

    ext[5, W]: extinction per km at unit profile value; asca[W]: aerosol scattering; g[W]: aerosol asymmetry;
    albedo[W]; strato_ext/strato_sca/strato_g[W]; mix[W] (Rayleigh depolarisation factor); solar[W]."""

    ext: np.ndarray
    asca: np.ndarray
    g: np.ndarray
    albedo: np.ndarray
    mix: np.ndarray
    solar: np.ndarray
    wavelengths: np.ndarray
    strato_ext: np.ndarray
    strato_sca: np.ndarray
    strato_g: np.ndarray
    lane_state: np.ndarray  # index of the state each lane belongs to
    states: list = field(default_factory=list)

    @property
    def W(self):
        return len(self.asca)

    @property
    def strato(self):
        return bool(np.any(self.strato_ext > 0))


def profile_column(spec, k, top):
    """Vertical column (km) of profile k of a spec from 0 to top: closed forms for the exponential and Gaussian kinds,
    Gauss-Legendre on 400 pieces for the others (piecewise-linear shapes are integrated exactly on the table nodes
    and at the triangle's corners, which fall on piece boundaries of the 0.5 km multiples used)."""
    kind = spec.kinds[k]
    if kind == EXP:
        return _column_exp(spec.a[k], top)
    if kind == GAUSS:
        return _column_gauss(spec.b[k], spec.a[k], top)
    x, w = np.polynomial.legendre.leggauss(8)
    edges = np.linspace(0.0, top, 401)
    h = ((edges[:-1] + edges[1:]) / 2)[:, None] + (np.diff(edges) / 2)[:, None] * x[None, :]
    vals = evaluate_profile(spec, k, h.ravel(), evaluate_profile(spec, 0, h.ravel())).reshape(h.shape)
    return float(np.sum(vals * w[None, :] * (np.diff(edges) / 2)[:, None]))


def lane_optics(state, wl):
    """Optics of one state at wavelengths wl (nm). Each component's extinction at unit profile is its column amount
    over the column of its profile shape."""
    wl = np.atleast_1d(np.asarray(wl, 'f8'))
    top = state.top_km
    spec = state.profiles
    ray = state.rayleigh_tau550 / profile_column(spec, 0, top) * (550.0 / wl) ** 4
    aer = state.aod550 / profile_column(spec, 1, top) * (550.0 / wl) ** state.angstrom
    oz_col = state.ozone_du * LOSCHMIDT_DU / (profile_column(spec, 2, top) * KM_CM)
    o3 = ozone_cross_section(wl) * oz_col * KM_CM
    no2_cols = profile_column(spec, 3, top)
    no2_n = state.no2_column / (no2_cols * KM_CM) if no2_cols > 0 else 0.0
    no2 = no2_cross_section(wl) * no2_n * KM_CM
    n_o2 = 0.20946 * 2.547e19  # O2 number density at the ground (cm^-3)
    o4 = o4_cross_section(wl) * n_o2 * n_o2 * KM_CM if state.o4 else np.zeros_like(wl)
    ext = np.stack([ray, aer, o3, no2, o4])
    if state.strato:
        col = profile_column(spec, 5, top)
        sext = state.sulfate_aod550 / col * (550.0 / wl) ** state.sulfate_angstrom
    else:
        sext = np.zeros_like(wl)
    W = len(wl)
    return Optics(
        ext=ext,
        asca=aer * state.aerosol_ssa,
        g=np.full(W, state.aerosol_g),
        albedo=np.full(W, state.albedo),
        mix=np.full(W, state.mix),
        solar=solar_irradiance(wl),
        wavelengths=wl,
        strato_ext=sext,
        strato_sca=sext * state.sulfate_ssa,
        strato_g=np.full(W, state.sulfate_g),
        lane_state=np.zeros(W, 'i4'),
        states=[state],
    )


PROFILE_KEYS = ('top_km', 'air_height_km', 'aerosol_height_km', 'ozone_center_km', 'ozone_sigma_km', 'no2_center_km',
                'no2_sigma_km', 'sulfate_center_km', 'sulfate_sigma_km', 'profile_spec')


def check_profile_shapes(state, base):
    """Raise E_BATCH_PROFILE unless `state` has the profile shapes and top of `base`: the geometry, sun table, ground
    pools and node geometry of a grid are built from one set of profiles, and each lane's extinction is normalised by
    its own state's columns, so a state of other shapes would get wrong optical depths."""
    for key in PROFILE_KEYS:
        if getattr(state, key) != getattr(base, key):
            raise ValueError(f'E_BATCH_PROFILE: state {state.name!r} differs from {base.name!r} in {key}')


def concat_optics(parts):
    """Pack the lanes of several states into one chunk. All parts must share the profile shapes (aerosol height,
    ozone/NO2/sulfate centres and widths, top), since the transport evaluates one set of profiles per sample."""
    base = parts[0].states[0]
    for p in parts:
        for s in p.states:
            check_profile_shapes(s, base)
    states = []
    lane_state = []
    for p in parts:
        lane_state.append(p.lane_state + len(states))
        states.extend(p.states)
    cat = lambda k: np.concatenate([getattr(p, k) for p in parts], axis=-1)
    return Optics(
        ext=cat('ext'), asca=cat('asca'), g=cat('g'), albedo=cat('albedo'), mix=cat('mix'), solar=cat('solar'),
        wavelengths=cat('wavelengths'), strato_ext=cat('strato_ext'), strato_sca=cat('strato_sca'),
        strato_g=cat('strato_g'), lane_state=np.concatenate(lane_state), states=states,
    )


def pad_optics(o, W):
    """Pad a chunk to W lanes by repeating its last lane (the reference pads the last 15-chunk the same way,
    `src/sosfuse2/jobs/run_state.py:592-599`)."""
    if o.W == W:
        return o
    if o.W > W:
        raise ValueError('E_WAVELENGTHS: chunk wider than the lane count')
    idx = np.r_[np.arange(o.W), np.full(W - o.W, o.W - 1)]
    take = lambda a: a[..., idx]
    return Optics(
        ext=take(o.ext), asca=take(o.asca), g=take(o.g), albedo=take(o.albedo), mix=take(o.mix), solar=take(o.solar),
        wavelengths=take(o.wavelengths), strato_ext=take(o.strato_ext), strato_sca=take(o.strato_sca),
        strato_g=take(o.strato_g), lane_state=take(o.lane_state), states=o.states,
    )


def hg_moments(g, L):
    """Legendre moments chi_l = (2l+1) g^l of a Henyey-Greenstein phase function (`docs/MODEL.md:236-242`), l = 0..L."""
    l = np.arange(L + 1)
    return (2 * l + 1) * np.asarray(g, 'f8')[..., None] ** l


def log_spaced_wavelengths(n, lo=400.0, hi=700.0):
    return np.exp(np.linspace(np.log(lo), np.log(hi), n))


def with_albedo(state, albedo):
    return replace(state, albedo=albedo)


def legacy_state(parameters, no2_profile, step_km=0.5, depolarisation=0.03, name='legacy'):
    """A state of the reference's legacy density (`src/sosfuse2/atmosphere/density.py:92,497-513`): air exp(-h/8),
    aerosol exp(-h/H) with H = parameters['scale_height_km'], the triangle ozone of half width 15 km about 25 km, NO2
    tabulated on a uniform grid of step_km from height 0 (no2_profile, maximum 1), O4 as the squared air profile, no
    sulfate layer, top 80 km. `parameters` are the reference state's (aod550, angstrom, ssa, scale_height_km, g,
    ozone_du, albedo); the coefficients per km come from the caller (optics_from_coefficients)."""
    no2 = np.asarray(no2_profile, 'f8')
    if no2.ndim != 1 or len(no2) < 2 or not np.all(np.isfinite(no2)) or np.any(no2 < 0):
        raise ValueError('E_PROFILE_TABLE: a finite non-negative NO2 profile of at least two nodes is required')
    if (len(no2) - 1) * step_km < 80.0:
        raise ValueError('E_PROFILE_TABLE: the NO2 profile must reach 80 km')
    rows = np.zeros((6, len(no2)))
    rows[3] = no2
    p = parameters
    H = float(p['scale_height_km'])
    spec = ProfileSpec(kinds=(EXP, EXP, TRIANGLE, TABLE, AIR_SQUARED, ZERO), a=(8.0, H, 15.0, float(step_km), 1.0, 1.0),
                       b=(0.0, 0.0, 25.0, 0.0, 0.0, 0.0), table=tuple(tuple(r) for r in rows))
    return State(name=name, top_km=80.0, depolarisation=depolarisation, aod550=float(p['aod550']),
                 aerosol_height_km=H, aerosol_g=float(p['g']), aerosol_ssa=float(p['ssa']),
                 angstrom=float(p['angstrom']), ozone_du=float(p['ozone_du']), albedo=float(p['albedo']),
                 profile_spec=spec)


def optics_from_coefficients(state, wl, ext, asca, solar):
    """Optics of one state from coefficients the caller computed: ext[5, W] per km at unit profile value (Rayleigh,
    aerosol, O3, NO2, O4 in the order of the reference's `extinction.coefficients`,
    `src/sosfuse2/atmosphere/extinction.py:25-33`), asca[W] the aerosol scattering per km and solar[W] the solar
    flux (`src/sosfuse2/solve/engine_core.py:86-92`). g, albedo and the depolarisation come from the state."""
    wl = np.atleast_1d(np.asarray(wl, 'f8'))
    W = len(wl)
    ext = np.asarray(ext, 'f8').reshape(NCOMP, W)
    asca = np.broadcast_to(np.asarray(asca, 'f8'), (W,)).copy()
    solar = np.broadcast_to(np.asarray(solar, 'f8'), (W,)).copy()
    if not (np.all(np.isfinite(ext)) and np.all(ext >= 0) and np.all(asca >= 0) and np.all(asca <= ext[1] * (1 + 1e-12))):
        raise ValueError('E_OPTICS: finite non-negative coefficients with asca <= aerosol extinction required')
    zero = np.zeros(W)
    return Optics(ext=ext, asca=asca, g=np.full(W, state.aerosol_g), albedo=np.full(W, state.albedo),
                  mix=np.full(W, state.mix), solar=solar, wavelengths=wl, strato_ext=zero, strato_sca=zero.copy(),
                  strato_g=np.full(W, state.sulfate_g), lane_state=np.zeros(W, 'i4'), states=[state])
