"""Autotuner of the launch shapes: the SH-records work-item size (nodes of one base cell per workgroup), the FP32
SH-records workgroup form (flow: double-buffered basis tiles with one barrier per round, two wavefronts per SIMD;
quad: the same rounds on four wavefronts per SIMD, each holding half of the K range, the next round's basis rows
fetched straight into LDS; staggered quad: the quad rounds with the two K halves forming the next tile before and
after their products; the stepwise form, seeds, basis and products separated by barriers, is never faster and is not
timed by default) and the transport's persistent workgroups per CU. None changes a result (every
record and every ray is computed the same way whatever the shape), so the tuner times one source evaluation per
candidate on the solver's own grid and lane count and keeps the fastest. Results are cached per (grid, L, lanes,
precision) in memory and, when a path is given, in a JSON file.
"""

import json
import os
import time

import torch

from .kernel import lane_count

NODE_CHUNKS = (256, 512, 1024, 2048)
BLOCKS_PER_CU = (4, 8, 16)
SH_TILES = (0, 2, 3)
_CACHE = {}


def key(solver):
    s, k = solver.s, solver.kernel
    return f'{s.geo.nh}x{s.geo.nc}-L{k.L}-G{lane_count(s.W)}-{"mixed" if k.mixed else "fp32"}-r{solver.geometry.bundle.rays}'


def _load(path):
    if path and os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def autotune(solver, node_chunks=NODE_CHUNKS, blocks_per_cu=BLOCKS_PER_CU, sh_tiles=SH_TILES, repeats=1, path=None,
             log=None):
    """Choose kernel.tuning for this solver's workload; returns (best, timings {candidate: seconds})."""
    k = solver.kernel
    kk = key(solver)
    stored = _load(path)
    if kk in _CACHE or kk in stored:
        best = _CACHE.get(kk) or stored[kk]
        k.tuning.update(best)
        return dict(best), {}
    x = solver.apply(solver.zeros())  # warm-up; builds the plans; a source with the real structure
    torch.cuda.synchronize()
    timings = {}
    tiles = sh_tiles if not k.mixed else (0,)
    for nc, b, st in ((nc, b, st) for nc in node_chunks for b in blocks_per_cu for st in tiles):
        k.tuning.update(node_chunk=nc, blocks_per_cu=b, sh_tile=st)
        solver.apply(x)  # applies the new item size outside the timed launch
        best_t = None
        for _ in range(repeats):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            solver.apply(x)
            e1.record()
            e1.synchronize()
            t = e0.elapsed_time(e1) * 1e-3
            best_t = t if best_t is None else min(best_t, t)
        timings[(nc, b, st)] = best_t
        if log:
            log(f'TUNE node_chunk={nc} blocks_per_cu={b} sh_tile={st}: {best_t * 1e3:.1f} ms')
    (nc, b, st), _ = min(timings.items(), key=lambda kv: kv[1])
    best = dict(node_chunk=nc, blocks_per_cu=b, sh_tile=st)
    k.tuning.update(best)
    _CACHE[kk] = best
    if path:
        stored[kk] = best
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        with open(path, 'w') as f:
            json.dump(stored, f, indent=1, sort_keys=True)
    return best, timings
