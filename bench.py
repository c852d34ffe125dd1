#!/usr/bin/env python3
"""Radiance tables per hour on one MI355X: the full plain FP32 table (150 levels x 75 chi, 30 wavelengths, L = 32,
25 solar elevations x 32 x 36 views, Stokes I/Q/U). One step = one table: the source iteration of all wavelengths to
relative tolerance 1e-5 plus the table query. Inputs (geometry, optics) are resident in HBM before the timed region.

Prints one JSON line (the bench contract) with a roofline entry for the dominant kernel (warm transport) and a CPU
baseline: the NumPy FP32 oracle timed on a bounded sample of the same work on this host.
"""

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_FP32 = 157.3e12  # MI355X FP32 (vector and matrix) FLOP/s
F_SAMPLE = 550  # FP32 flops per (quadrature node, wavelength) of the warm transport, counted from the kernel source
F_VISIT_PER_K = 2  # flops per (SH corner refill, wavelength, coefficient): one FMA


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(s, scope, evaluations, seconds_budget=20.0):
    """Time the FP32 oracle (NumPy, one core) on a bounded sample of cells of the same workload and extrapolate to
    one table (evaluations x warm+ground transport and projection, plus first light)."""
    import numpy as np

    from oracle.transport import PAIRS4, project, transport
    from sosfuse2_amd import geometry

    os.environ.setdefault('OMP_NUM_THREADS', '1')
    c = s.constants(s.geo.path_n)
    rng = np.random.default_rng(0)
    field = (np.abs(rng.normal(size=(s.geo.ns, s.rows, s.W))) * 1e-4).astype('f4')
    cells = list(range(0, s.geo.ns, max(1, s.geo.ns // 24)))
    total_rays = sum(len(p['d']) for p in scope.packs)
    done_rays = 0
    warm_t = first_t = 0.0
    t_start = time.perf_counter()
    for cell in cells:
        g = scope.packs[cell]
        t = time.perf_counter()
        v = transport(g, c, field, False, np.zeros((s.geo.nc, s.W)), PAIRS4)
        project(g, v, c)
        warm_t += time.perf_counter() - t
        t = time.perf_counter()
        v = transport(g, c, field, True, None, PAIRS4)
        project(g, v, c)
        first_t += time.perf_counter() - t
        done_rays += len(g['d'])
        if time.perf_counter() - t_start > seconds_budget:
            break
    scale = total_rays / done_rays
    table_s = scale * (first_t + evaluations * warm_t * (1 + 1.0 / s.geo.nh))
    return dict(value=3600.0 / table_s, unit='tables/hour', cores=1, kind='port',
                sample=f'FP32 NumPy oracle (flush-to-zero), warm and first-light transport + projection of '
                       f'{done_rays} rays in {len(cells[:len(cells)])} cells sampled across the grid, all {s.W} '
                       f'wavelengths, extrapolated to {total_rays} rays x {evaluations} evaluations per table '
                       f'({table_s:.0f} s per table)')


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=2)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--config', default='config1', choices=['config0', 'config1'])
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--mixed', action='store_true', help='FP16 SH plane (mixed precision variant)')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f'bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}); '
              'the work shards by state and wavelength chunk with no exchange, see DESIGN.md', file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from sosfuse2_amd import api, atmosphere

    cfg = dict(api.CONFIG1 if args.config == 'config1' else api.CONFIG0)
    state = atmosphere.config1_state() if args.config == 'config1' else atmosphere.config0_state()
    wl, hs, el, az = api.axes(cfg)
    t0 = time.perf_counter()
    s, scope = api.create_configured(None, cfg, wl, 'hip', state, mixed=args.mixed)
    solver = scope.solver(s, log=None)
    k = s.kernel
    k.plan(solver.geometry.bundle, k.constants(s, solver.geometry.bundle.nq))
    k.plan(solver.geometry.ground_bundle, k.constants(s, solver.geometry.ground_bundle.nq))
    torch.cuda.synchronize()
    prep = time.perf_counter() - t0
    log(f'prepared in {prep:.1f} s: {s.geo.ns} cells, {solver.geometry.bundle.rays} rays, '
        f'{solver.geometry.bundle.samples} nodes per evaluation, W={s.W}')
    tol = float(cfg['anderson_tol'])

    def step():
        # a table of a new state: first light, the source iteration and the query (geometry and plans are shared)
        sv = scope.solver(s, log=log)
        x, info = sv.solve(60, tol)
        table, _ = api.make_table(s, x, hs, el, az, log=log)
        return info, table

    for i in range(args.warmup):
        t = time.perf_counter()
        info, _ = step()
        torch.cuda.synchronize()
        log(f'warmup {i}: {time.perf_counter() - t:.2f} s, {info["evaluations"]} evaluations')
    k.timed = []
    torch.cuda.synchronize()
    t = time.perf_counter()
    evals = []
    for i in range(args.steps):
        info, table = step()
        evals.append(info['evaluations'])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t
    timed, k.timed = k.timed, None
    per_step = elapsed / args.steps
    warm = [a.elapsed_time(b) * 1e-3 for name, a, b in timed if name == 'warm']
    kinds = {}
    for name, a, b in timed:
        kinds.setdefault(name, []).append(a.elapsed_time(b) * 1e-3)
    log('launch seconds per step: ' + ', '.join(f'{n} {sum(v) / args.steps:.3f} ({len(v) // args.steps})'
                                                  for n, v in kinds.items()))
    # algorithmic work of one warm launch (fixed by the geometry): nodes and SH refills from the device counters
    k.counters = torch.zeros(2, dtype=torch.int64, device=k.device)
    x = solver.zeros()
    dc = k.constants(s, solver.geometry.bundle.nq)
    plane = k.plane(s, x, dc)
    bnd = torch.zeros((s.geo.nc, dc.G), dtype=torch.float32, device=k.device)
    k.radiance(solver.geometry.bundle, s, x, False, bnd, 4, plane)
    torch.cuda.synchronize()
    nodes, refills = (int(v) for v in k.counters.cpu())
    k.counters = None
    flops = (nodes * F_SAMPLE + refills * F_VISIT_PER_K * k.K) * s.W
    t_warm = sum(warm) / max(1, len(warm))
    achieved = flops / t_warm / 1e12
    roofline = dict(bound='mfma', achieved=round(achieved, 3), peak=PEAK_FP32 / 1e12, unit='TFLOP/s',
                    frac=round(achieved / (PEAK_FP32 / 1e12), 4), traffic=None)
    log(f'warm transport: {t_warm * 1e3:.2f} ms per launch, nodes {nodes}, refills {refills}, '
        f'{flops / 1e12:.3f} TFLOP -> {achieved:.2f} TFLOP/s')
    out = dict(
        metric='radiance tables/hour/GPU at <=1e-4 rel err vs FP64; learned+batched vs plain',
        value=round(3600.0 / per_step, 2),
        unit='tables/hour',
        n_gpus=1,
        steps=args.steps,
        warmup=args.warmup,
        ms_per_step=round(per_step * 1e3, 2),
        higher_is_better=True,
        scaling='weak',
        vs_baseline=None,
        dtype='fp16-sh' if args.mixed else 'f32',
        data='synthetic',
        config=dict(workload=f'{args.config}: plain FP32 HIP, {s.geo.nh} levels x {s.geo.nc} chi, '
                             f'{s.W} wavelengths, L={k.L}, nmu={s.geo.nmu} nphi={s.geo.nphi} path_n={s.geo.path_n}, '
                             f'{len(hs)} solar elevations x {len(el)} x {len(az)} views, I/Q/U, tol {tol:g}',
                    evaluations=evals, rays=solver.geometry.bundle.rays, prepare_seconds=round(prep, 1)),
        roofline=roofline,
    )
    if not args.no_cpu:
        out['cpu_baseline'] = cpu_baseline(s, scope, int(np.mean(evals)))
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
