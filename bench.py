#!/usr/bin/env python3
"""Radiance tables per hour on one MI355X: the full plain FP32 table (150 levels x 75 chi, 30 wavelengths, L = 32,
25 solar elevations x 32 x 36 views, Stokes I/Q/U). One step = one table of a state: its optics and device constants,
the source iteration of all wavelengths to relative tolerance 1e-5 and the table query. The grid's geometry (bundles,
visit plans, node geometry, basis tables) is resident in HBM before the timed region.

Prints one JSON line (the bench contract) with a roofline entry for the dominant kernel (the SH contraction of the
warm pass, sh_records) and a CPU baseline: the NumPy FP32 oracle timed on a bounded sample of the same work on this
host. --mixed runs the mixed-precision form (split-FP16 SH contraction on v_mfma_f32_16x16x32_f16).
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

PEAK_FP32 = 157.3e12  # MI355X dense FP32 MFMA FLOP/s
PEAK_FP16 = 2.5e15  # MI355X dense FP16 MFMA FLOP/s
PEAK_HBM = 8.0e12  # MI355X HBM3E bytes/s
F_SAMPLE = 550  # FP32 flops per (quadrature node, wavelength) of the warm transport, counted from the kernel source
COPY_BYTES = 1 << 30  # the calibration copy reads and writes this many bytes


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(kernel, config, mixed):
    """HBM bytes per launch of the dominant kernel, from the committed rocprofv3 counter summary of this workload
    (FETCH_SIZE + WRITE_SIZE, scaled by the same counters on a copy of known size); None when absent."""
    name = f'{kernel}_{config}_{"fp16" if mixed else "fp32"}_traffic.txt'
    path = ROOT / 'profiles' / name
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith('bytes_per_launch:'):
            return float(line.split(':', 1)[1])
    return None


def calibration_copy():
    """A plain device copy of COPY_BYTES (read and write) that counter runs use to scale FETCH_SIZE / WRITE_SIZE."""
    import torch

    n = COPY_BYTES // 4
    x = torch.ones(n, dtype=torch.float32, device='cuda')
    y = torch.empty_like(x)
    torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    log(f'calibration copy: {COPY_BYTES} bytes read and written')
    del x, y


def batched_variant(scope, n, el, az, tol, seed=2026):
    """Tables/hour of the batched form: n states of the batched distribution (uneven wavelength and elevation
    counts) packed into 32-lane chunks, solved and queried; the plain form's rate of the same states solved one by
    one is estimated from its lanes (each state alone is one chunk)."""
    import torch

    from sosfuse2_amd import batch

    items = batch.items_from_seed(n, seed, top_km=120.0)
    chunks = batch.pack(items, 32)
    torch.cuda.synchronize()
    t = time.perf_counter()
    res = batch.solve(scope, items, el, az, tolerance=tol, max_eval=60, lanes=32, log=log)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t
    return dict(states=n, seed=seed, chunks=len(chunks), lanes=sum(len(it.wavelengths) for it in items),
                elevations=sum(len(it.elevations) for it in items), seconds=round(elapsed, 2),
                tables_per_hour=round(3600.0 * n / elapsed, 2),
                evaluations=[r[1]['evaluations'] for r in res])


def learned_variant(scope, ntrain, ntest, el, az, tol):
    """The learned warm start in the batched form: train the network on the solves of ntrain states, then solve
    ntest held-out states cold and warm-started (both to the same tolerance) and compare evaluations and time."""
    import numpy as np
    import torch

    from sosfuse2_amd import atmosphere, batch, learned

    t0 = time.perf_counter()
    train_items = batch.items_from_seed(ntrain, 7, top_km=120.0)
    rng = np.random.default_rng(0)
    Xs, Ts = [], []

    for ch in batch.pack(train_items, 32):
        optics = atmosphere.concat_optics([atmosphere.lane_optics(train_items[i].state, train_items[i].wavelengths)
                                           for i in ch.members])
        s = scope.context(optics)
        sv = scope.solver(s, log=None)
        x, info = sv.solve(60, tol)
        X, T = learned.samples(s, sv, x, cells_per_lane=2048, rng=rng)
        Xs.append(X)
        Ts.append(T)
        log(f'LEARN data chunk {ch.members}: {info["evaluations"]} evaluations')
    X, T = torch.cat(Xs), torch.cat(Ts)
    gen = time.perf_counter() - t0
    net = learned.Network(learned.sizes_for(s.rows), seed=0)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    hist = learned.train(net, X, T, 1500, batch=8192, lr=1e-3, log=log)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t1
    del X, T, Xs, Ts
    test_items = batch.items_from_seed(ntest, 1234, top_km=120.0)
    torch.cuda.synchronize()
    t = time.perf_counter()
    cold = batch.solve(scope, test_items, el, az, tolerance=tol, max_eval=60, lanes=32)
    torch.cuda.synchronize()
    t_cold = time.perf_counter() - t
    out = dict(train_states=ntrain, test_states=ntest, data_seconds=round(gen, 1), train_seconds=round(train_s, 1),
               loss_first=hist[0], loss_last=hist[-1],
               cold=dict(seconds=round(t_cold, 2), tables_per_hour=round(3600.0 * ntest / t_cold, 2),
                         evaluations=[r[1]['evaluations'] for r in cold]))
    # the warm start over all source rows, and over the tensor and isotropic rows only
    for key, rows in (('warm', 'all'), ('warm_low_orders', 'low')):
        t = time.perf_counter()
        warm = batch.solve(scope, test_items, el, az, tolerance=tol, max_eval=60, lanes=32,
                           x0=lambda s, sv: learned.warm_start(net, s, sv, rows=rows))
        torch.cuda.synchronize()
        t_warm = time.perf_counter() - t
        worst = max(float(np.max(np.abs(w[0][..., 0] - c[0][..., 0]) / np.abs(c[0][..., 0])))
                    for w, c in zip(warm, cold))
        out[key] = dict(seconds=round(t_warm, 2), tables_per_hour=round(3600.0 * ntest / t_warm, 2),
                        evaluations=[r[1]['evaluations'] for r in warm], max_I_difference_vs_cold=worst)
    return out


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
    ap.add_argument('--no-tune', action='store_true', help='keep the default launch shapes (no autotuning)')
    ap.add_argument('--mixed', action='store_true', help='headline in the mixed-precision form (split-FP16 SH)')
    ap.add_argument('--variants', default='mixed,batched',
                    help='forms measured beside the headline: mixed, batched (comma list, or "none")')
    ap.add_argument('--batched-states', type=int, default=8, help='states of the batched sample')
    ap.add_argument('--learned', type=str, default='',
                    help='TRAIN,TEST: train the warm-start network on TRAIN batched states, compare TEST states '
                         'warm-started against cold (adds minutes)')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f'bench.py runs on one GPU: there is no multi-GPU path (asked for --gpus {args.gpus}); '
              'the work would shard by state and wavelength chunk, one process per GPU, with no exchange',
              file=sys.stderr)
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
    calibration_copy()
    tuning = None
    if not args.no_tune:
        from sosfuse2_amd import tune

        t = time.perf_counter()
        best, timings = tune.autotune(solver, node_chunks=(512, 1024, 2048), blocks_per_cu=(8, 16), log=log)
        tuning = dict(best, seconds=round(time.perf_counter() - t, 1), candidates=len(timings))
        log(f'autotuned in {tuning["seconds"]} s: {best}')
    log(f'prepared in {prep:.1f} s: {s.geo.ns} cells, {solver.geometry.bundle.rays} rays, '
        f'{solver.geometry.bundle.samples} nodes per evaluation, W={s.W}')
    tol = float(cfg['anderson_tol'])

    def step():
        # a table of a new state: its lanes' optics and device constants (ground pool values on the device), first
        # light, the source iteration and the query; the grid's geometry, bundles, visit plans, node geometry and
        # basis tables are shared by every state
        sc = scope.context(atmosphere.lane_optics(state, wl))
        sv = scope.solver(sc, log=log)
        x, info = sv.solve(60, tol)
        table, _ = api.make_table(sc, x, hs, el, az, log=log)
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
    launch_s = {n: round(sum(v) / args.steps, 4) for n, v in kinds.items()}
    # the dominant kernel is the SH contraction of the warm pass: 2 K W flops per (node, corner) visit
    sh = [x for x in kinds.get('sh_warm', [])]
    t_sh = sum(sh) / max(1, len(sh))
    plan = k.plan(solver.geometry.bundle, k.constants(s, solver.geometry.bundle.nq))
    visits = int(plan.vis.numel())
    flops = 2.0 * visits * k.K * s.W
    peak = PEAK_FP16 if args.mixed else PEAK_FP32
    achieved = flops / t_sh
    traffic = measured_traffic('sh_records', args.config, args.mixed)
    roofline = dict(bound='mfma', achieved=round(achieved / 1e12, 3), peak=peak / 1e12, unit='TFLOP/s',
                    frac=round(achieved / peak, 4), traffic=traffic)
    log(f'SH records (warm): {t_sh * 1e3:.2f} ms per launch, {visits} visits x K {k.K} x W {s.W} -> '
        f'{flops / 1e12:.3f} TFLOP, {achieved / 1e12:.2f} TFLOP/s ({"FP16" if args.mixed else "FP32"} MFMA)')
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
        dtype='f32+fp16-mfma' if args.mixed else 'f32',
        data='synthetic',
        config=dict(workload=f'{args.config}: {"mixed FP16 SH" if args.mixed else "plain FP32"} HIP, {s.geo.nh} levels x {s.geo.nc} chi, '
                             f'{s.W} wavelengths, L={k.L}, nmu={s.geo.nmu} nphi={s.geo.nphi} path_n={s.geo.path_n}, '
                             f'{len(hs)} solar elevations x {len(el)} x {len(az)} views, I/Q/U, tol {tol:g}',
                    evaluations=evals, rays=solver.geometry.bundle.rays, prepare_seconds=round(prep, 1),
                    tuning=tuning),
        roofline=roofline,
        launch_seconds_per_step=launch_s,
    )
    variants = [v for v in args.variants.split(',') if v and v != 'none']
    if args.mixed or 'mixed' in variants:
        # the same state's table in the other form (same kernel, plans and state): per-entry error of I, Q, U of the
        # mixed table against the FP32 one
        k.mixed = not args.mixed
        torch.cuda.synchronize()
        t = time.perf_counter()
        info_o, table_o = step()
        torch.cuda.synchronize()
        t_other = time.perf_counter() - t
        k.mixed = args.mixed
        t16, t32 = (table, table_o) if args.mixed else (table_o, table)
        I32 = np.abs(t32[..., 0])
        d = np.abs(t16 - t32) / I32[..., None]
        err = dict(max_I=float(d[..., 0].max()), rms_I=float(np.sqrt(np.mean(d[..., 0] ** 2))),
                   max_QU=float(d[..., 1:].max()))
        if args.mixed:
            out['error_vs_fp32'] = dict(err, evaluations_fp32=info_o['evaluations'])
        else:
            out['mixed'] = dict(tables_per_hour=round(3600.0 / t_other, 2), evaluations=info_o['evaluations'],
                                error_vs_fp32=err, form='split-FP16 SH contraction (v_mfma_f32_16x16x32_f16)')
        log(f'mixed vs FP32 table: {err}')
    if 'batched' in variants and args.batched_states > 0:
        out['batched'] = batched_variant(scope, args.batched_states, el, az, tol)
    if args.learned:
        ntrain, ntest = (int(v) for v in args.learned.split(','))
        out['learned'] = learned_variant(scope, ntrain, ntest, el, az, tol)
    if not args.no_cpu:
        out['cpu_baseline'] = cpu_baseline(s, scope, int(np.mean(evals)))
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
