"""Batched solves: many atmosphere states with uneven wavelength and solar-elevation counts.

The lanes of a launch are (state, wavelength) pairs, so states that share the grid and the profile shapes are packed
into chunks of up to 32 lanes (first-fit decreasing on the wavelength count, at most four states per chunk so the
first-light ground records keep one phase group per state). Each chunk is one source iteration: every lane keeps
its own Anderson history and convergence flag (per-wavelength histories, `src/sosfuse2/solve/source_iteration.py:
432-443`), and the chunk stops when its last lane converges. The table of a chunk is queried once over the union of
its states' solar elevations and split back per state. The reference runs one state per job and cuts its wavelength
list into 15-lane chunks (`src/sosfuse2/jobs/run_state.py:592-599,839`); packing states fills the lanes instead.
"""

import dataclasses
import time

import numpy as np

from . import api, atmosphere


@dataclasses.dataclass
class Item:
    state: atmosphere.State
    wavelengths: np.ndarray
    elevations: list  # solar elevations (degrees) of this state's table


@dataclasses.dataclass
class Chunk:
    members: list  # item indices
    lanes: int


def items_from_seed(n, seed, top_km=120.0, all_elevations=None, max_wavelengths=30, max_elevations=25):
    """The batched workload: n seeded states (atmosphere.sample_states) with their uneven wavelength counts
    (log-spaced over 400-700 nm) and solar elevations (the first counts of +6 ... -18 in 1 degree steps)."""
    grid = list(all_elevations) if all_elevations is not None else [6.0 - i for i in range(25)]
    out = []
    for st, nw, ne in atmosphere.sample_states(n, seed, top_km, max_wavelengths, max_elevations):
        wl = atmosphere.log_spaced_wavelengths(nw) if nw > 1 else np.array([550.0])
        out.append(Item(st, np.asarray(wl, float), grid[: min(ne, len(grid))]))
    return out


def pack(items, lanes=32, max_states=4):
    """First-fit decreasing by wavelength count into chunks of at most `lanes` lanes and `max_states` states."""
    order = sorted(range(len(items)), key=lambda i: (-len(items[i].wavelengths), i))
    chunks = []
    for i in order:
        w = len(items[i].wavelengths)
        if w > lanes:
            raise ValueError(f'E_WAVELENGTHS: state {i} has {w} wavelengths, more than {lanes} lanes')
        for ch in chunks:
            if ch.lanes + w <= lanes and len(ch.members) < max_states:
                ch.members.append(i)
                ch.lanes += w
                break
        else:
            chunks.append(Chunk([i], w))
    return chunks


def solve_chunk(scope, items, el, az, s, tolerance=1e-5, max_eval=60, x0=None, graph=True):
    """Solve one packed chunk (its context `s` holds the members' lanes in order) and query its table over the union
    of the members' solar elevations: per member (table [h_s, el, az, W_i, 3] I/Q/U, info)."""
    solver = scope.solver(s, log=None, graph=graph)
    start = x0(s, solver) if x0 is not None else None
    x, info = solver.solve(max_eval, tolerance, x0=start)
    hs = sorted({h for it in items for h in it.elevations}, reverse=True)
    table, _ = api.make_table(s, x, hs, el, az, log=None)
    out, lane = [], 0
    lanes = sum(len(it.wavelengths) for it in items)
    for it in items:
        w = len(it.wavelengths)
        rows = [hs.index(h) for h in it.elevations]
        out.append((table[rows][..., lane : lane + w, :], dict(info, lanes=lanes)))
        lane += w
    return out


def solve(scope, items, el, az, tolerance=1e-5, max_eval=60, lanes=32, log=None, x0=None, graph=True):
    """Solve every item and return per item (table [h_s, el, az, W_i, 3] I/Q/U, info). `x0(s)` may give a warm
    start for a chunk context (the learned warm start); the iteration still runs to `tolerance`. graph: each chunk's
    evaluation runs as one captured HIP graph, replayed every iteration."""
    for it in items:
        atmosphere.check_profile_shapes(it.state, scope.state)
    out = [None] * len(items)
    for ch in pack(items, lanes):
        t0 = time.perf_counter()
        optics = atmosphere.concat_optics([atmosphere.lane_optics(items[i].state, items[i].wavelengths)
                                           for i in ch.members])
        s = scope.context(optics)
        res = solve_chunk(scope, [items[i] for i in ch.members], el, az, s, tolerance, max_eval, x0, graph)
        for i, r in zip(ch.members, res):
            out[i] = (r[0], dict(r[1], chunk=ch.members))
        if log:
            log(f'BATCH chunk states={ch.members} lanes={ch.lanes} evaluations={res[0][1]["evaluations"]} '
                f'seconds={time.perf_counter() - t0:.2f}')
    return out
