"""Write tests/golden/config0_oracle.npz: config0 solved by the oracle, the checker of the GPU tests.

config0: 60 levels to 100 km on the 10-node chi grid, 550 nm, L = 16, Anderson depth 5, tolerance 1e-5, the table
at solar elevations -3 and -6 over 4 x 6 views with Stokes I, Q, U. Stored: the FP32-FTZ oracle's converged source,
evaluation count and table, and the FP64 twin's table (solved to 1e-7). Run from the repository root:
    python tests/golden/make_config0.py [workers]
"""

import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))


def main():
    from types import SimpleNamespace

    import numpy as np

    from oracle.solve import OracleSolver, make_pool
    from sosfuse2_amd import api, atmosphere

    workers = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    cfg = dict(api.CONFIG0, workers=workers)
    wl, hs, el, az = api.axes(cfg)
    st = atmosphere.config0_state()
    scope = api.Scope(dict(cfg, sun_scale=1), st, SimpleNamespace(L=int(cfg['L'])))
    s = scope.context(atmosphere.lane_optics(st, wl))
    out = dict(wavelengths=wl, solar_elevation=hs, view_elevation=el, azimuth=az)
    for name, dtype, tol in (('fp32', np.float32, float(cfg['anderson_tol'])), ('fp64', np.float64, 1e-7)):
        pool = make_pool(s, scope.packs, dtype, workers)
        t = time.perf_counter()
        o = OracleSolver(s, dtype, depth=int(cfg['anderson_depth']), packs=scope.packs, pool=pool, workers=workers)
        x, info = o.solve(80, tol)
        table = o.table(x, hs, el, az)
        pool.close()
        print(f'{name}: {info["evaluations"]} evaluations, {time.perf_counter() - t:.1f} s', flush=True)
        out[f'{name}_table'] = table
        out[f'{name}_evaluations'] = np.int64(info['evaluations'])
        if name == 'fp32':
            out['fp32_source'] = x
    np.savez_compressed(Path(__file__).with_name('config0_oracle.npz'), **out)


if __name__ == '__main__':
    main()
