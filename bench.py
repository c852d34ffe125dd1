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
    """HBM bytes per launch of a kernel, from the committed rocprofv3 counter summary of this workload (FETCH_SIZE +
    WRITE_SIZE, scaled by the same counters on a copy of known size); None when absent."""
    name = f'{kernel}_{config}_{"fp16" if mixed else "fp32"}_traffic.txt'
    path = ROOT / 'profiles' / name
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith('bytes_per_launch:'):
            return float(line.split(':', 1)[1])
    return None


def angular_pairs(plan, nc):
    """P of the transport's SH work: the (ray, corner cell) refills of the reference's per-ray SH cache
    (numpy_reference.py:579-600, counted as `angular_pairs` at :597) after its adjacent-stencil aliases
    (reuse_corners, :106-129: when the base cell moves by one row and/or column the cache entries of the corners the
    old and new stencils share are copied, not refilled). So a corner cell of a node is refilled exactly when it is
    not a corner of the ray's previous node (or the node is the ray's first): P = the number of maximal runs of
    consecutive nodes of one ray that share a corner cell. The four corners of base cell b are b, b + 1, b + nc,
    b + nc + 1 (nc chi nodes per row; base cells never sit in the last column or row)."""
    import torch

    n = plan.nodes
    cell = plan.node_cell[:n]
    starts = plan.node_off[:-1][plan.node_off[:-1] < plan.node_off[1:]]
    first = torch.zeros(n, dtype=torch.bool, device=cell.device)
    first[starts] = True
    first[0] = True
    prev = torch.empty_like(cell)
    prev[0] = cell[0]
    prev[1:] = cell[:-1]
    total = 0
    for off in (0, 1, nc, nc + 1):
        d = cell + off - prev
        shared = (d == 0) | (d == 1) | (d == nc) | (d == nc + 1)
        total += int((first | ~shared).sum().item())
        del d, shared
    return total


def consumer_bytes(plan, rays, cells, G):
    """Bytes the warm consumer (transport_geom_kernel) must move per launch given its inputs: the SH records
    [4 nodes][G] and the node geometry [nodes][32] (FP32), the tensor rows of the source [cells][4][G], the per-ray
    constants (start, direction, edges, end columns: about 120 bytes) and the radiance written [rays][G][4]."""
    n = plan.nodes
    return 4.0 * (4 * n * G + 32 * n + 4 * cells * G + 4 * rays * G) + 120.0 * rays


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


def normalised_tables(items):
    """Table equivalents of a set of states: lanes x solar elevations over the full table's 30 x 25."""
    return sum(len(it.wavelengths) * len(it.elevations) for it in items) / (30 * 25)


def batched_run(scope, items, el, az, tol, mixed=False, x0=None, graph=True):
    """Solve and query every state in packed chunks (one captured HIP graph per chunk's evaluation)."""
    import torch

    from sosfuse2_amd import batch

    k = scope.kernel
    k.mixed, was = mixed, k.mixed
    try:
        torch.cuda.synchronize()
        t = time.perf_counter()
        res = batch.solve(scope, items, el, az, tolerance=tol, max_eval=60, lanes=32, log=log, x0=x0, graph=graph)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t
    finally:
        k.mixed = was
    n = len(items)
    return dict(states=n, seconds=round(elapsed, 2), tables_per_hour=round(3600.0 * n / elapsed, 2),
                normalised_tables_per_hour=round(3600.0 * normalised_tables(items) / elapsed, 2),
                evaluations=[r[1]['evaluations'] for r in res]), res


def batched_variant(scope, n, el, az, tol, seed=2026):
    """The batched form: n states of the batched distribution (uneven wavelength and elevation counts) packed into
    32-lane chunks, each chunk's evaluation replayed as a captured HIP graph, FP32. The normalised rate counts table
    equivalents (lanes x elevations over 30 x 25), comparable with the plain table rate."""
    from sosfuse2_amd import batch

    items = batch.items_from_seed(n, seed, top_km=120.0)
    out, _ = batched_run(scope, items, el, az, tol)
    return dict(out, seed=seed, chunks=len(batch.pack(items, 32)), lanes=sum(len(it.wavelengths) for it in items),
                elevations=sum(len(it.elevations) for it in items),
                table_equivalents=round(normalised_tables(items), 3)), items


def train_network(scope, ntrain, tol, seed=7, steps=1500):
    """Train the warm-start network on the converged sources of ntrain states of the batched distribution."""
    import numpy as np
    import torch

    from sosfuse2_amd import atmosphere, batch, learned

    t0 = time.perf_counter()
    train_items = batch.items_from_seed(ntrain, seed, top_km=120.0)
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
    X, T = torch.cat(Xs), torch.cat(Ts)
    gen = time.perf_counter() - t0
    net = learned.Network(learned.sizes_for(s.rows), seed=0)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    hist = learned.train(net, X, T, steps, batch=8192, lr=1e-3, log=None)
    torch.cuda.synchronize()
    info = dict(train_states=ntrain, samples=int(X.shape[0]), data_seconds=round(gen, 1),
                train_seconds=round(time.perf_counter() - t1, 1), loss_first=hist[0], loss_last=hist[-1])
    return net, info


def learned_variant(scope, items, el, az, tol, ntrain, cold_mixed):
    """The learned + mixed-precision + batched form on the batched states: the network (trained on other states)
    seeds each chunk, the SH contraction runs in split FP16, every chunk replays a captured graph; the iteration
    still runs to the same tolerance. Compared with the same states cold in mixed precision (the start's gain) and
    against the plain table rate."""
    import numpy as np

    from sosfuse2_amd import learned

    net, info = train_network(scope, ntrain, tol)
    out, res = batched_run(scope, items, el, az, tol, mixed=True,
                           x0=lambda s, sv: learned.warm_start(net, s, sv, rows='low'))
    worst = max(float(np.max(np.abs(w[0][..., 0] - c[0][..., 0]) / np.abs(c[0][..., 0])))
                for w, c in zip(res, cold_mixed))
    return dict(out, training=info, max_I_difference_vs_cold=worst,
                evaluations_total=int(sum(out['evaluations'])),
                evaluations_total_cold=int(sum(r[1]['evaluations'] for r in cold_mixed)))


_CPU = {}


def _cpu_cell(job):
    """Warm (with a boundary) and first-light transport + projection of every `step`-th ray of one config1 cell by
    the FP32 oracle, in a forked worker that inherited the host context; returns (rays, warm s, first-light s)."""
    import numpy as np

    from oracle.transport import PAIRS4, project, transport

    cell, step = job
    c, packs, field, nc = _CPU['c1'], _CPU['packs1'], _CPU['field1'], _CPU['nc1']
    g = packs[cell]
    n = len(g['d'])
    g = {k: (v[::step] if isinstance(v, np.ndarray) and v.shape[:1] == (n,) else v) for k, v in g.items()}
    t = time.perf_counter()
    v = transport(g, c, field, False, np.zeros((nc, c['W'])), PAIRS4)
    project(g, v, c)
    warm = time.perf_counter() - t
    t = time.perf_counter()
    v = transport(g, c, field, True, None, PAIRS4)
    project(g, v, c)
    return len(g['d']), warm, time.perf_counter() - t


def _ftz_active(_):
    import numpy as np

    from oracle.arith import flush_to_zero

    tiny = np.float32(np.finfo(np.float32).tiny)
    with flush_to_zero():
        return bool(tiny * np.float32(0.5) == 0)


class CpuBaseline:
    """The FP32 NumPy oracle on this host's cores: a process pool forked from this process before it starts the GPU
    runtime (the workers inherit the host contexts and never touch the GPU), one worker per core of the affinity
    mask (at most 16). Two measurements: config1 (the bench workload) on a bounded sample of cells, extrapolated to
    a table, and config0 end to end (solve to 1e-5 at Anderson depth 5 and its 2 x 4 x 6 table)."""

    def __init__(self, scope1, ctx1, cfg0, state0):
        import multiprocessing as mp
        from types import SimpleNamespace

        import numpy as np

        from oracle import solve as osolve
        from sosfuse2_amd import api, atmosphere

        self.cores = max(1, min(16, len(os.sched_getaffinity(0))))
        rng = np.random.default_rng(0)
        ns = ctx1.geo.ns
        _CPU.update(c1=ctx1.constants(ctx1.geo.path_n), packs1=scope1.packs, nc1=ctx1.geo.nc,
                    field1=(np.abs(rng.normal(size=(ns, ctx1.rows, ctx1.W))) * 1e-4).astype('f4'))
        self.scope1, self.ctx1 = scope1, ctx1
        wl0, self.hs0, self.el0, self.az0 = api.axes(cfg0)
        self.cfg0 = cfg0
        scope0 = api.Scope(dict(cfg0, sun_scale=1, workers=1), state0, SimpleNamespace(L=int(cfg0['L'])))
        self.ctx0 = scope0.context(atmosphere.lane_optics(state0, wl0))
        self.packs0 = scope0.packs
        osolve._worker_init(self.ctx0.constants(self.ctx0.geo.path_n), self.ctx0.constants(self.ctx0.geo.query_n),
                            self.packs0)
        self.pool = mp.get_context('fork').Pool(self.cores)
        self.pool.map(_ftz_active, range(self.cores))  # the workers exist before the GPU runtime starts

    def run(self, evaluations, per_core=1, step=1):
        import numpy as np

        from oracle.solve import OracleSolver

        ctx1 = self.ctx1
        pick = np.unique(np.linspace(0, ctx1.geo.ns - 1, per_core * self.cores).astype(int)).tolist()
        total_rays = sum(len(p['d']) for p in self.scope1.packs)
        t = time.perf_counter()
        parts = self.pool.map(_cpu_cell, [(c, step) for c in pick], chunksize=1)
        wall = time.perf_counter() - t
        rays = sum(p[0] for p in parts)
        warm, first = sum(p[1] for p in parts), sum(p[2] for p in parts)
        # one table = first light + evaluations x (warm pass + ground pass, the ground pass being the first row of
        # cells); CPU seconds scale with rays, spread over the pool's cores
        table_cpu = total_rays / rays * (first + evaluations * warm * (1 + 1.0 / ctx1.geo.nh))
        table_s = table_cpu / self.cores
        ftz = all(self.pool.map(_ftz_active, range(self.cores)))
        t = time.perf_counter()
        o = OracleSolver(self.ctx0, np.float32, depth=int(self.cfg0['anderson_depth']), packs=self.packs0,
                         pool=self.pool, workers=self.cores)
        x, info = o.solve(60, float(self.cfg0['anderson_tol']))
        table0 = o.table(x, self.hs0, self.el0, self.az0)
        t0 = time.perf_counter() - t
        self.pool.close()
        threads = os.environ.get('OMP_NUM_THREADS', 'unset')
        return dict(
            value=3600.0 / table_s, unit='tables/hour', cores=self.cores, kind='port',
            sample=f'FP32 NumPy oracle (flush-to-zero {"active" if ftz else "INACTIVE"}) in {self.cores} worker '
                   f'processes (one per core of the affinity mask, OMP_NUM_THREADS={threads}): warm and first-light '
                   f'transport + projection of {rays} rays ({"all" if step == 1 else f"every {step}th"} ray of '
                   f'{len(pick)} cells spread over the config1 grid, one cell per worker), all {ctx1.W} wavelengths, {wall:.1f} s wall ({warm + first:.0f} CPU s), extrapolated to {total_rays} '
                   f'rays x {evaluations} evaluations per table ({table_cpu:.0f} CPU s, {table_s:.0f} s per table '
                   f'on {self.cores} cores)',
            config0_end_to_end=dict(seconds=round(t0, 2), tables_per_hour=round(3600.0 / t0, 2),
                                    evaluations=info['evaluations'], cores=self.cores,
                                    what=f'config0 solved by the FP32 oracle to {self.cfg0["anderson_tol"]:g} '
                                         f'(Anderson depth {self.cfg0["anderson_depth"]}) plus its '
                                         f'{len(self.hs0)} x {len(self.el0)} x {len(self.az0)} table, cells and '
                                         f'observer packs spread over the pool'),
        )


def config0_variant(tol_default=None):
    """config0 on the GPU end to end (60 levels, 550 nm, L = 16, depth 5, tolerance 1e-5, its 2 x 4 x 6 table): the
    workload the CPU baseline also solves end to end."""
    import torch

    from sosfuse2_amd import api, atmosphere

    cfg = dict(api.CONFIG0)
    wl, hs, el, az = api.axes(cfg)
    s, scope = api.create_configured(None, cfg, wl, 'hip', atmosphere.config0_state())
    tol = float(cfg['anderson_tol'])
    scope.solver(s, log=None).solve(60, tol)  # warm-up: plans and constants
    torch.cuda.synchronize()
    t = time.perf_counter()
    reps = 3
    for _ in range(reps):
        sc = scope.context(atmosphere.lane_optics(atmosphere.config0_state(), wl))
        x, info = scope.solver(sc, log=None).solve(60, tol)
        api.make_table(sc, x, hs, el, az, log=None)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t) / reps
    return dict(seconds=round(sec, 4), tables_per_hour=round(3600.0 / sec, 1), evaluations=info['evaluations'],
                cells=s.geo.ns)


PRODUCTION = dict(nmu=12, nphi=16, path_n=12, sun_scale=4)  # the reference's full-grid quadrature


def production_variant(pscope, kernel, state, tol, headline_table, release):
    """config1 at the reference's production quadrature (nmu 12, nphi 16, path_n 12, sun table scale 4): one table
    of the state (device geometry, first light, the solve to the tolerance, the 25-elevation I/Q/U table), the warm
    and first-light passes run in parts of at most Kernel.part_nodes quadrature nodes with one visit plan resident
    at a time. Reported beside the headline with the table's difference from the headline's (coarser) quadrature."""
    import numpy as np
    import torch

    from sosfuse2_amd import api, atmosphere

    release()  # the headline's resident plans
    torch.cuda.empty_cache()
    torch.cuda.reset_peak_memory_stats()
    cfg = pscope.cfg
    wl, hs, el, az = api.axes(cfg)
    pscope.kernel = kernel
    t = time.perf_counter()
    pscope.finish_on_device()
    s = pscope.context(atmosphere.lane_optics(state, wl))
    sv = pscope.solver(s, log=log)
    db = sv.geometry.bundle
    parts = kernel.parts(db, kernel.constants(s, db.nq))
    torch.cuda.synchronize()
    geo_s = time.perf_counter() - t
    log(f'production quadrature: {db.rays} rays, {db.nodes_total} nodes in {len(parts)} parts, geometry {geo_s:.1f} s')
    t = time.perf_counter()
    x, info = sv.solve(60, tol)
    table, _ = api.make_table(s, x, hs, el, az, log=log)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t
    I, I0 = table[..., 0], headline_table[..., 0]
    return dict(workload=f'config1 at nmu={cfg["nmu"]} nphi={cfg["nphi"]} path_n={cfg["path_n"]} '
                         f'sun_scale={cfg["sun_scale"]} (the reference full-grid quadrature), plain FP32, one table',
                seconds=round(sec, 2), tables_per_hour=round(3600.0 / sec, 2), evaluations=info['evaluations'],
                rays=db.rays, nodes=int(db.nodes_total), parts=len(parts), part_nodes=kernel.part_nodes,
                geometry_seconds=round(geo_s, 1), peak_device_gb=round(torch.cuda.max_memory_allocated() / 1e9, 1),
                max_I_difference_vs_headline_quadrature=float(np.max(np.abs(I - I0) / np.abs(I0))),
                rms_I_difference_vs_headline_quadrature=float(np.sqrt(np.mean(((I - I0) / I0) ** 2))))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=2)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--config', default='config1', choices=['config0', 'config1'])
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-tune', action='store_true', help='keep the default launch shapes (no autotuning)')
    ap.add_argument('--mixed', action='store_true', help='headline in the mixed-precision form (split-FP16 SH)')
    ap.add_argument('--variants', default='mixed,batched,learned,config0,production',
                    help='forms measured beside the headline: mixed, batched, learned (learned + mixed + batched, '
                         'needs batched), config0, production (config1 at the reference full-grid quadrature: '
                         'one table, about 30 s) (comma list, or "none")')
    ap.add_argument('--batched-states', type=int, default=64, help='states of the batched sample')
    ap.add_argument('--learned-train', type=int, default=16, help='training states of the warm-start network')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f'bench.py runs on one GPU: there is no multi-GPU path (asked for --gpus {args.gpus}); '
              'the work would shard by state and wavelength chunk, one process per GPU, with no exchange',
              file=sys.stderr)
        sys.exit(2)
    from types import SimpleNamespace

    import numpy as np
    import torch

    from sosfuse2_amd import api, atmosphere
    from sosfuse2_amd.kernel import Kernel

    cfg = dict(api.CONFIG1 if args.config == 'config1' else api.CONFIG0)
    state = atmosphere.config1_state() if args.config == 'config1' else atmosphere.config0_state()
    wl, hs, el, az = api.axes(cfg)
    t0 = time.perf_counter()
    # host preparation (grid, sun table, ground pools, cell packs) and the CPU baseline's worker processes come
    # before anything starts the GPU runtime
    scope = api.Scope(dict(cfg, sun_scale=1), state, SimpleNamespace(L=int(cfg['L'])))
    variants = [v for v in args.variants.split(',') if v and v != 'none']
    pscope = None
    if 'production' in variants and args.config == 'config1':
        # host part of the production geometry (ground-pool directions); the columns, quadrature, packs and
        # bundles follow on the device
        pscope = api.Scope(dict(cfg, **PRODUCTION), state, None, device_columns=True, device_geometry=True)
    cpu = None
    if not args.no_cpu and args.config == 'config1':
        cpu = CpuBaseline(scope, scope.context(atmosphere.lane_optics(state, wl)), dict(api.CONFIG0),
                          atmosphere.config0_state())
    scope.kernel = Kernel(backend='hip', L=int(cfg['L']), mixed=args.mixed)
    s = scope.context(atmosphere.lane_optics(state, wl))
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
    kinds = {}
    for name, a, b in timed:
        kinds.setdefault(name, []).append(a.elapsed_time(b) * 1e-3)
    log('launch seconds per step: ' + ', '.join(f'{n} {sum(v) / args.steps:.3f} ({len(v) // args.steps})'
                                                  for n, v in kinds.items()))
    launch_s = {n: round(sum(v) / args.steps, 4) for n, v in kinds.items()}
    # the dominant kernel is the SH contraction of the warm pass; its algorithmic work is the reference's:
    # 2 K W flops per (ray, corner cell) refill of the per-ray SH cache after its corner aliases (angular_pairs)
    sh = kinds.get('sh_warm', [])
    t_sh = sum(sh) / max(1, len(sh))
    db = solver.geometry.bundle
    plan = k.plan(db, k.constants(s, db.nq))
    pairs = angular_pairs(plan, s.geo.nc)
    visits = int(plan.visits)
    flops = 2.0 * pairs * k.K * s.W
    peak = PEAK_FP16 if args.mixed else PEAK_FP32
    achieved = flops / t_sh
    traffic = measured_traffic('sh_records', args.config, args.mixed)
    roofline = dict(bound='mfma', achieved=round(achieved / 1e12, 3), peak=peak / 1e12, unit='TFLOP/s',
                    frac=round(achieved / peak, 4), traffic=traffic,
                    kernel=f'sh_records ({"split-FP16" if args.mixed else "FP32"} MFMA), warm pass',
                    work=f'2 K W per (ray, corner cell) refill: {pairs} refills (angular_pairs) x K {k.K} x W {s.W}; '
                         f'the node-centric kernel computes {visits} (node, corner) visits, '
                         f'{visits / pairs:.3f}x the refills',
                    ms_per_launch=round(t_sh * 1e3, 3))
    log(f'SH records (warm): {t_sh * 1e3:.2f} ms per launch, {pairs} refills ({visits} visits) x K {k.K} x W {s.W} '
        f'-> {flops / 1e12:.3f} TFLOP, {achieved / 1e12:.2f} TFLOP/s ({"FP16" if args.mixed else "FP32"} MFMA)')
    warm = kinds.get('warm', [])
    t_c = sum(warm) / max(1, len(warm))
    G = k.constants(s, db.nq).G
    cbytes = consumer_bytes(plan, db.rays, s.geo.ns, G)
    ctraffic = measured_traffic('transport_geom', args.config, args.mixed)
    transport_roofline = dict(
        kernel='transport_geom_kernel (warm consumer)', bound='hbm', achieved=round(cbytes / t_c / 1e9, 1),
        peak=PEAK_HBM / 1e9, unit='GB/s', frac=round(cbytes / t_c / PEAK_HBM, 4), traffic=ctraffic,
        measured=round(ctraffic / t_c / 1e9, 1) if ctraffic else None, ms_per_launch=round(t_c * 1e3, 3),
        bytes=f'{cbytes / 1e9:.2f} GB per launch: records 4 x {plan.nodes} nodes x {G} lanes, node geometry 32 '
              f'floats per node, tensor rows, ray constants and radiance out')
    log(f'warm consumer: {t_c * 1e3:.2f} ms per launch, {cbytes / 1e9:.2f} GB -> {cbytes / t_c / 1e9:.0f} GB/s')
    del plan  # the counts are taken; the bundle alone holds its plan (released before the production table)
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
                    evaluations=evals, rays=db.rays, prepare_seconds=round(prep, 1),
                    tuning=tuning),
        roofline=roofline,
        transport_roofline=transport_roofline,
        launch_seconds_per_step=launch_s,
    )
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
    if 'batched' in variants and args.batched_states > 0 and args.config == 'config1':
        out['batched'], items = batched_variant(scope, args.batched_states, el, az, tol)
        out['batched']['vs_plain'] = round(out['batched']['normalised_tables_per_hour'] / out['value'], 3)
        log(f'batched: {out["batched"]}')
        if 'learned' in variants:
            cold, cold_res = batched_run(scope, items, el, az, tol, mixed=True)
            out['batched_mixed'] = dict(cold, vs_plain=round(cold['normalised_tables_per_hour'] / out['value'], 3))
            lm = learned_variant(scope, items, el, az, tol, args.learned_train, cold_res)
            lm['vs_plain'] = round(lm['normalised_tables_per_hour'] / out['value'], 3)
            out['learned_mixed_batched'] = lm
            log(f'learned + mixed + batched: {lm}')
    if 'config0' in variants:
        out['config0'] = config0_variant()
    if pscope is not None:
        def release():
            solver.geometry.bundle.plan = None
            solver.geometry.ground_bundle.plan = None
            # the earlier tables' query bundles (one per elevation set of the batched chunks) and their plans
            from sosfuse2_amd.solver import drop_observer_bundles

            drop_observer_bundles(k)

        out['production'] = production_variant(pscope, k, state, tol, table, release)
        out['production']['vs_headline'] = round(out['production']['tables_per_hour'] / out['value'], 4)
        log(f'production quadrature: {out["production"]}')
    if cpu is not None:
        out['cpu_baseline'] = cpu.run(int(round(np.mean(evals))))
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
