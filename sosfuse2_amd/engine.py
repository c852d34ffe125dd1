"""Per-state engine context: grid, optics, sun table, ground pools, the constants of the transport and the observer
queries (the role of the reference's `src/sosfuse2/solve/engine_core.py:77-238` Solver object, `s`).

A `Context` is what the reference's `create_configured` returns as `s`: the solver `Solver` iterates its source and
`make_table` queries it. Geometry that depends only on the grid (cell packs, bundles, columns) is shared by every
state with the same grid and profile shapes; the per-lane optics are cheap to swap.
"""

import math
import time

import numpy as np

from . import atmosphere, geometry, layout


class Context:
    """One solve chunk: a grid, the optics of its W lanes, and the per-grid geometry."""

    def __init__(self, grid, optics, kernel, sun=None, pools=None, observer_km=0.1, profile_state=None,
                 monitors=None, directions=None, record_monitors=None):
        self.geo = grid
        self.optics = optics
        self.kernel = kernel
        self.state = profile_state or optics.states[0]
        self.observer_km = observer_km
        self.W = optics.W
        self.wl = optics.wavelengths
        self.solar = optics.solar
        self.strato = optics.strato
        t = time.perf_counter()
        self.sun = sun if sun is not None else geometry.sun_table(self.state, grid.domain)
        self.pools = pools if pools is not None else geometry.ground_pools(self.state, grid, self.sun)
        self.prepare_seconds = time.perf_counter() - t
        stop, record = grid.domain.monitors()
        # stop monitors decide convergence; record monitors are evaluated and logged only
        self.monitors = tuple(stop if monitors is None else monitors)
        self.record_monitors = tuple(record if record_monitors is None else record_monitors)
        # Observer boxes of the convergence monitor (`solve/engine.py:435-443`): 7 directions x 4 x 4 GL boxes.
        self.directions = directions or [(2, 0), (10, 0), (30, 0), (10, 180), (30, 180), (10, 90), (89, 0)]
        self._constants = {}
        self._base = {}
        self._gv = None
        self._padded = None
        self._gm = None

    @property
    def L(self):
        return self.kernel.L

    @property
    def K(self):
        return layout.count(self.kernel.L)

    @property
    def rows(self):
        return layout.source_rows(self.kernel.L, self.strato)

    def with_optics(self, optics):
        """The same grid and geometry with other lanes (another state or wavelength chunk)."""
        return Context(self.geo, optics, self.kernel, self.sun, self.pools, self.observer_km, self.state,
                       self.monitors, self.directions, self.record_monitors)

    # -------------------------------------------------------------------------------------------- constants
    def pool_values(self):
        """gv[P, W] of the ground pools for this context's lanes (computed once)."""
        if self._gv is None:
            self._gv = self.pools.values(self.optics)
        return self._gv

    def constants(self, nq, dtype=np.float32, padded=True):
        """Transport constants (the reference's Kernel.setup, `src/sosfuse2/solve/source_solver.py:1243-1339`).
        padded: include the pool arrays gm (per-cell tensor moments) and the per-cell zero-padded gd, gv (read by the
        oracle; the device path computes its own from the lane table, sos_ground_pools).
        Everything but the Gauss-Legendre nodes is shared by every node count."""
        key = (nq, np.dtype(dtype).str, padded)
        c = self._constants.get(key)
        if c is not None:
            return c
        base = self._base.get(np.dtype(dtype).str)
        if base is None:
            base = self._base[np.dtype(dtype).str] = self._base_constants(dtype)
        u, wu = geometry.gl(nq)
        c = dict(base, u=u, wu=wu)
        if padded:
            F = np.dtype(dtype).type
            if self._padded is None:
                self._padded = self.pools.padded(self.optics, self.geo.ns, self.pool_values())
                self._gm = self.pools.tensor(self.optics, self.geo.ns, self.pool_values())
            gd, gv = self._padded
            c.update(gd=gd.astype(F), gv=gv.astype(F), gm=self._gm.astype(F))
        self._constants[key] = c
        return c

    def _base_constants(self, dtype):
        F = np.dtype(dtype).type
        g, o, st = self.geo, self.optics, self.state
        prof = atmosphere.profiles(st, g.h)
        return dict(
            F=F,
            W=o.W,
            L=self.kernel.L,
            K=self.K,
            top=float(g.domain.top_km),
            h=g.h.astype(F),
            chi=g.chi.astype(F),
            hr=prof[:, 0].astype(F),
            ha=prof[:, 1].astype(F),
            ext=o.ext.astype(F),
            asca=o.asca.astype(F),
            g=o.g.astype(F),
            mix=o.mix.astype(F),
            albedo=o.albedo.astype(F),
            strato=bool(self.strato),
            strato_ext=o.strato_ext.astype(F),
            strato_sca=o.strato_sca.astype(F),
            strato_g=o.strato_g.astype(F),
            strato_density=atmosphere.strato_profile(st, g.h).astype(F),
            sun_h=self.sun.h.astype(F),
            sun_col=self.sun.col.astype(F),
            profile=dict(air_H=st.air_height_km, aer_H=st.aerosol_height_km, o3_c=st.ozone_center_km,
                         o3_s=st.ozone_sigma_km, no2_c=st.no2_center_km, no2_s=st.no2_sigma_km,
                         st_c=st.sulfate_center_km, st_s=st.sulfate_sigma_km, spec=st.profiles),
            combine=True,
        )

    def eigen(self):
        """Per-row composition factors [rows, W] that turn incident moments into the source
        (`src/sosfuse2/solve/source_iteration.py:176-208`): 3/(8 pi) ext_R for the tensor, rf ext_R for rows 4-9,
        g^l asca for the aerosol plane, g_s^l sca_s for the sulfate plane."""
        o = self.optics
        L, K = self.kernel.L, self.K
        lv = layout.degrees(L)
        out = np.zeros((self.rows, o.W))
        out[:4] = (3 / (8 * math.pi)) * o.ext[0][None, :]
        out[4:10] = layout.rayleigh_factors()[:, None] * o.ext[0][None, :]
        out[10 : 10 + K] = o.g[None, :] ** lv[:, None] * o.asca[None, :]
        if self.strato:
            out[10 + K :] = o.strato_g[None, :] ** lv[:, None] * o.strato_sca[None, :]
        return out

    # ---------------------------------------------------------------------------------------------- queries
    def query_packs(self, h_s, el, az, chunk=64):
        """Observer rays for one solar elevation, in packs of `chunk` rays (`engine_core.py:208-226`)."""
        ee, aa = np.broadcast_arrays(np.asarray(el, 'f8'), np.asarray(az, 'f8'))
        e, z = ee.ravel(), aa.ravel()
        out = []
        for j in range(0, len(e), chunk):
            out.append(geometry.observer_pack(self.geo, h_s, e[j : j + chunk], z[j : j + chunk], self.observer_km))
        return out

    def box_directions(self):
        """(el, az, weight) of the 7 x 16 observer box rays (`engine_core.py:228-238`)."""
        u, w = np.polynomial.legendre.leggauss(4)
        ww = (w[:, None] * w[None, :] / 4).ravel()
        el = np.array([e + de for e, z in self.directions for de in u for dz in u])
        az = np.array([z + dz for e, z in self.directions for de in u for dz in u])
        return el, az, ww
