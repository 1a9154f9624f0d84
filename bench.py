"""Benchmark: agent-steps/s of the per-timestep agent update on one MI355X.

Workload: 20 000 Circular agents in a 200 x 100 m room with four 1.2 m exits,
potential-gradient navigation on a 0.1 m grid, block-list cell 0.6 m, eight
wall segments, adaptive dt in [0.001, 0.01], fluctuation on (the standard
logic tree).  A step is one MultiAgentSimulation.update(): binning, pair
forces, walls, adjusting, navigation, fluctuation, integration.  Inputs are
resident in HBM before timing; steps are replayed from a HIP graph.

Prints one JSON line.  `python bench.py --gpus N --steps K --warmup W`.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ('agent-steps/s @20k Circular (GPU fp64/fp32 vs host numba); adjoint step time')
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
KERNELS = ('bin_insert', 'cell_gather', 'forces', 'integrate')


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pair_tally(arr, cell, segments):
    """Algorithmic work of one step's force pass on state `arr`, by the
    per-pair operation tally (10 / 80 / 126 ops per candidate pair that is
    rejected / social / social+contact; 22 / 52 per agent-wall pair without /
    with contact), counting each unordered pair once as the reference does."""
    from oracle import oracle as O
    pairs = O.iter_pairs(arr['position'], cell)
    p = arr['position']
    r = arr['radius']
    i, j = pairs[:, 0], pairs[:, 1]
    h = np.hypot(*(p[i] - p[j]).T) - (r[i] + r[j])
    n_rej = int(np.sum(h >= 3.0))
    n_con = int(np.sum(h < 0.0))
    n_soc = len(h) - n_rej - n_con
    seg = np.concatenate([segments['p0'], segments['p1']], 1) if len(segments) else np.zeros((0, 4))
    from crowddynamics_amd.examples import _seg_dist
    if len(seg):
        dw = _seg_dist(p[:, 0], p[:, 1], seg) - r[:, None]
        w_con = int(np.sum(dw < 0))
        w_all = dw.size
    else:
        w_con = w_all = 0
    flops = 10 * n_rej + 80 * n_soc + 126 * n_con + 22 * (w_all - w_con) + 52 * w_con
    return dict(candidate_pairs=len(h), social_pairs=n_soc, contact_pairs=n_con,
                wall_pairs=w_all, wall_contacts=w_con, flops=flops)


def binning_bytes(n, cells):
    """24 B per agent (read position, write sorted index) + 16 B per cell
    (count and offset)."""
    return 24 * n + 16 * cells


def integrate_bytes(n, real_bytes=8):
    """Circular adjust+integrate traffic: reads x, v, f, f_prev, e0 and m,
    tau_adj, v0; writes x, v, f_prev (152 B/agent in fp64)."""
    return n * (16 + 4 * 2 * real_bytes + 3 * 8 + 16 + 2 * 2 * real_bytes)


def cpu_baseline(scn, seconds, threads=None):
    """Oracle on the host: single-thread restatement of the reference's
    algorithm (threads=None) or its OpenMP owner-computes variant."""
    from oracle import oracle as O
    arr = scn.agents.copy()
    nv = scn.nav
    sc = O.Scenario(cell_size=scn.cell_size, dt_min=scn.dt_min, dt_max=scn.dt_max,
                    seed=scn.seed, segments=scn.segments, nav_maps=nv['maps'],
                    nav_mode=nv['mode'], nav_origin=(nv['minx'], nv['miny']),
                    nav_step=nv['step'])
    O.run(arr, sc, 2, threads=threads)  # warm-up
    steps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        O.run(arr, sc, 5, first_step=2 + steps, threads=threads)
        steps += 5
    el = time.perf_counter() - t0
    return scn.n * steps / el, steps, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=500)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--precision', default='f64', choices=('f64', 'f32'))
    ap.add_argument('--graph-steps', type=int, default=50)
    ap.add_argument('--cpu-seconds', type=float, default=10.0)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--n', type=int, default=20000)
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU: the per-step update runs on a single MI355X '
              '(multi-GPU sharding is described in DESIGN.md, not launched).', flush=True)
        sys.exit(2)

    import torch
    from crowddynamics_amd import _lib
    from crowddynamics_amd import examples as E
    from crowddynamics_amd.stepper import Stepper

    log('building workload: room n=%d' % args.n)
    scn = E.room(n=args.n)
    results = {}
    kernel_ms = {}
    stream = torch.cuda.Stream()
    for prec in ('f64', 'f32'):
        with torch.cuda.stream(stream):
            st = Stepper(scn.agents.copy(), segments=scn.segments, nav=scn.nav,
                         cell_size=scn.cell_size, dt_min=scn.dt_min, dt_max=scn.dt_max,
                         seed=scn.seed, precision=prec, bounds=scn.bounds)
            st.step(args.warmup)
            G = max(1, min(args.graph_steps, args.steps))
            graph = st.capture(G)
            reps, rem = divmod(args.steps, G)
            stream.synchronize()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                graph.launch()
            st.step(rem)
            stream.synchronize()
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            state = st.check()
            out = st.download(scn.agents.copy())
            # per-kernel GPU time, HIP events on the launch stream
            nt = 50
            ms = st.timed(nt)
            kernel_ms[prec] = {k: ms[i] / nt for i, k in enumerate(KERNELS)}
            results[prec] = dict(value=scn.n * args.steps / el, ms_per_step=1e3 * el / args.steps,
                                 time_tot=state.time_tot, cells=int(state.shape[0] * state.shape[1]),
                                 out=out)
            del graph
        log('%s: %.3e agent-steps/s, %.4f ms/step, kernels %s' % (
            prec, results[prec]['value'], results[prec]['ms_per_step'],
            {k: round(v * 1e3, 2) for k, v in kernel_ms[prec].items()}))

    main_p = args.precision
    km = kernel_ms[main_p]
    cells = results[main_p]['cells']
    t0_tally = pair_tally(scn.agents, scn.cell_size, scn.segments)
    t1_tally = pair_tally(results[main_p]['out'], scn.cell_size, scn.segments)
    flops_step = 0.5 * (t0_tally['flops'] + t1_tally['flops'])
    cand_step = 0.5 * (t0_tally['candidate_pairs'] + t1_tally['candidate_pairs'])
    rb = 8 if main_p == 'f64' else 4
    per_kernel = {}
    bin_ms = km['bin_insert'] + km['cell_gather']
    per_kernel['binning'] = dict(bound='hbm', unit='GB/s', ms=bin_ms,
                                 achieved=binning_bytes(scn.n, cells) / (bin_ms * 1e-3) / 1e9,
                                 peak=HBM_PEAK_GBS)
    per_kernel['integrate'] = dict(bound='hbm', unit='GB/s', ms=km['integrate'],
                                   achieved=integrate_bytes(scn.n, rb) / (km['integrate'] * 1e-3) / 1e9,
                                   peak=HBM_PEAK_GBS)
    per_kernel['forces'] = dict(bound='fp64-valu', unit='TFLOP/s', ms=km['forces'],
                                achieved=flops_step / (km['forces'] * 1e-3) / 1e12,
                                peak=FP64_PEAK_TFLOPS,
                                candidate_pairs_per_s=cand_step / (km['forces'] * 1e-3))
    for v in per_kernel.values():
        v['frac'] = v['achieved'] / v['peak']
    dom = max(KERNELS, key=lambda k: km[k])
    if dom == 'forces':
        rk = per_kernel['forces']
        traffic = None
    elif dom == 'integrate':
        rk = per_kernel['integrate']
        traffic = None
    else:
        rk = dict(per_kernel['binning'])
        traffic = None
    roofline = dict(kernel=dom, bound=rk['bound'], achieved=rk['achieved'], peak=rk['peak'],
                    unit=rk['unit'], frac=rk['frac'], traffic=traffic)

    cpu = None
    if not args.no_cpu:
        log('cpu baseline (oracle, 1 thread, %.0f s sample)' % args.cpu_seconds)
        v1, s1, e1 = cpu_baseline(scn, args.cpu_seconds)
        from oracle import oracle as O
        nth = min(O.max_threads(), 16)
        log('cpu baseline (oracle OpenMP, %d threads)' % nth)
        vn, sn, en = cpu_baseline(scn, max(2.0, args.cpu_seconds / 3), threads=nth)
        cpu = dict(value=v1, unit='agent-steps/s', cores=1, kind='port',
                   sample='%d steps of the same 20k-agent room from its initial state, '
                          'fp64, %.1f s' % (s1, e1),
                   all_core=dict(value=vn, cores=nth, steps=sn))

    r = results[main_p]
    line = {
        'metric': METRIC,
        'value': r['value'],
        'unit': 'agent-steps/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': r['ms_per_step'],
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64' if main_p == 'f64' else 'f32 (f64 positions and dt)',
        'data': 'synthetic (seeded placement, closed-form exit potential)',
        'config': {'workload': 'room: 20k Circular agents, 200x100 m, 4 exits of 1.2 m, '
                               'cell 0.6 m, 0.1 m potential grid, adaptive dt',
                   'agents': scn.n, 'cell_size': scn.cell_size, 'precision': main_p,
                   'graph_steps': args.graph_steps},
        'roofline': roofline,
        'cpu_baseline': cpu,
        'fp32_value': results['f32']['value'] if main_p == 'f64' else results['f64']['value'],
        'kernel_ms': {p: {k: v for k, v in d.items()} for p, d in kernel_ms.items()},
        'rooflines': per_kernel,
        'pair_tally': t0_tally,
    }
    if cpu:
        line['speedup_vs_cpu_1core'] = r['value'] / cpu['value']
        line['speedup_vs_cpu_allcore'] = r['value'] / cpu['all_core']['value']
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
