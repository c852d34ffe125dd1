"""Write tests/golden/walkthrough.json: the inputs and the frozen outputs of the reference's reduced walkthrough state
(`src/sosfuse2/jobs/walkthrough.py:183-196` example_state, run by `src/sosfuse2/tests/test_offline_cpu.py:24-57`).

The state is clear_34 of `src/sosfuse2/assets/states_v2.json` with O4 off, the stratospheric layer disabled and the
legacy density (exp8 air, triangle ozone, NO2 of `assets/no2_afglus_profile.json`). The script reads the reference's
JSON assets as data and restates the arithmetic that turns them into the solver's per-km coefficients at 550 nm:

- Rayleigh: `src/sosfuse2/atmosphere/rayleigh.py:13-35` (Bodhaine 1999 eqs. 18, 19, 5, 6, 23, 22 at the vacuum
  wavelength, scaled to 0.013558 (550/450)^4 at 450 nm), vacuum wavelength by six fixed-point steps of Ciddor's
  standard-air index (`src/sosfuse2/atmosphere/air_wavelength.py:22-27`);
- aerosol: aod550 over the column of exp(-h/H) from 0 to 80 km (`src/sosfuse2/solve/engine.py:414-416`);
- ozone: the spectra.json cross section times ozone_du 2.687e20 / 15 (`src/sosfuse2/atmosphere/extinction.py:29`);
- NO2: the 220 K cross section of `assets/no2_vandaele1998_fine.json` interpolated onto the spectra.json grid and at
  550 nm (`src/sosfuse2/atmosphere/minor_absorbers.py:116-134`), times 2.687e20 no2_du / H_f
  (`extinction.py:14-18`), with no2_du the 0.207 default (`src/sosfuse2/atmosphere/components.py:143-147`) and H_f
  the trapezoid of the maximum-normalised profile over 0-80 km (`src/sosfuse2/atmosphere/density.py:384-443`);
- O4 off: zero; solar flux: spectra.json 'solar' at 550 nm (`src/sosfuse2/solve/engine_core.py:92`).

The expected table values are the frozen ones of `test_offline_cpu.py:51-56` (table[:, ::3, ::5, 0]) and the sunward
column with the observed record of `docs/WALKTHROUGH.md` (5 significant digits).

Run in a container that holds the reference: python tests/golden/make_walkthrough.py
"""

import json
from pathlib import Path

import numpy as np

ASSETS = Path('/root/reference/src/sosfuse2/assets')
OUT = Path(__file__).with_name('walkthrough.json')

CIDDOR = (238.0185, 5792105.0, 57.362, 167917.0)


def standard_n(vac_nm, co2=450.0):
    """Ciddor 1996 eq. 1, standard dry air, at the vacuum wavelength (nm)."""
    k0, k1, k2, k3 = CIDDOR
    s2 = 1.0 / (np.asarray(vac_nm, 'f8') * 1e-3) ** 2
    return 1.0 + (k1 / (k0 - s2) + k3 / (k2 - s2)) * 1e-8 * (1.0 + 0.534e-6 * (co2 - 450.0))


def vacuum_nm(air_nm):
    air = np.asarray(air_nm, 'f8')
    vac = air.copy()
    for _ in range(6):
        vac = air * standard_n(vac)
    return vac


def bodhaine_sigma(vac_nm):
    """Bodhaine et al. 1999 cross section (cm^2): refractive index of 300 ppm air scaled to 360 ppm CO2 (eqs. 18,
    19), King factor of N2, O2, Ar, CO2 (eqs. 5, 6, 23), sigma = 24 pi^3 (n^2-1)^2 / (lambda^4 Ns^2 (n^2+2)^2) F."""
    u = np.asarray(vac_nm, 'f8') / 1000.0
    n = 1 + (8060.51 + 2480990 / (132.274 - u ** -2) + 17455.7 / (39.32957 - u ** -2)) * 1e-8 * (
        1 + 0.54 * (360e-6 - 0.0003))
    fn2 = 1.034 + 3.17e-4 / u ** 2
    fo2 = 1.096 + 1.385e-3 / u ** 2 + 1.448e-4 / u ** 4
    king = (78.084 * fn2 + 20.946 * fo2 + 0.934 + 0.036 * 1.15) / (78.084 + 20.946 + 0.934 + 0.036)
    lam_cm = np.asarray(vac_nm, 'f8') * 1e-7
    return 24 * np.pi ** 3 * (n * n - 1) ** 2 / (lam_cm ** 4 * 2.546899e19 ** 2 * (n * n + 2) ** 2) * king


def rayleigh_per_km(air_nm):
    return 0.013558 * (550.0 / 450.0) ** 4 * bodhaine_sigma(vacuum_nm(air_nm)) / bodhaine_sigma(vacuum_nm(450.0))


def main():
    states = json.loads((ASSETS / 'states_v2.json').read_text())
    pool = states['states'] if isinstance(states, dict) and 'states' in states else states
    state = next(s for s in pool if s['id'] == 'clear_34')
    p = dict(state['parameters'])
    spectra = json.loads((ASSETS / 'spectra.json').read_text())
    fine = json.loads((ASSETS / 'no2_vandaele1998_fine.json').read_text())
    profile = json.loads((ASSETS / 'no2_afglus_profile.json').read_text())
    wl = 550.0
    sw = np.asarray(spectra['wavelength_nm'], 'f8')

    z = np.asarray(profile['altitude_km'], 'f8')
    v = np.asarray(profile['no2_number_density_m3'], 'f8')
    assert np.array_equal(z, np.arange(161) * 0.5)
    f = v / v.max()
    h_f = float(np.trapezoid(f, z))
    no2_du = 0.207
    no2_sigma = float(np.interp(wl, sw, np.interp(sw, fine['wavelength_nm'], fine['no2_220K'])))

    H = p['scale_height_km']
    ext = [
        float(rayleigh_per_km(wl)),
        p['aod550'] / (H * -np.expm1(-80.0 / H)) * (550.0 / wl) ** p['angstrom'],
        float(np.interp(wl, sw, spectra['ozone'])) * p['ozone_du'] * 2.687e20 / 15,
        no2_sigma * 2.687e20 * no2_du / h_f,
        0.0,
    ]
    doc = dict(
        source='clear_34 of the reference assets/states_v2.json, O4 off, stratosphere disabled, exp8/triangle/keep',
        parameters=p,
        no2_du=no2_du,
        depolarisation=0.03,
        wavelength_nm=[wl],
        ext_per_km=ext,
        asca_per_km=ext[1] * p['ssa'],
        solar=float(np.interp(wl, sw, spectra['solar'])),
        no2_profile_step_km=0.5,
        no2_profile=f.tolist(),
        no2_H_f_km=h_f,
        settings=dict(h=[0.0, 0.1, 8.0, 80.0], chi_deg=[0.0, 60.0, 85.0, 90.0, 93.0, 96.0, 105.0, 110.0, 140.0, 180.0],
                      nmu=1, nphi=2, path_n=4, L=16, anderson_tol=1e-3, anderson_depth=4, observer_km=0.1, query_n=6,
                      sun_scale=4, top_km=80.0, solar_elevation_deg=[-3.0, -6.0], view_shape=[4, 6], max_eval=64),
        expected=dict(
            corners=[[[3.054727e-4, 2.808648e-4], [3.163053e-3, 1.111142e-3]],
                     [[1.492353e-5, 1.283569e-5], [2.165351e-4, 4.395965e-5]]],
            corners_rtol=1e-3,
            sunward=[[3.0547e-04, 5.0828e-04, 1.0735e-03, 3.1631e-03], [1.4924e-05, 2.7430e-05, 7.7560e-05, 2.1654e-04]],
            minimum=1.284e-05, maximum=0.003163, rays=982, source_coefficients=153, evaluations=12,
        ),
    )
    OUT.write_text(json.dumps(doc, indent=1) + '\n')
    print('wrote', OUT, 'ext', ext)


if __name__ == '__main__':
    main()
