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
# the step's kernels and their slots in Stepper.timed() (slot 2, a separate
# integration kernel, is empty: the force kernel integrates)
KERNELS = {'step_begin': 0, 'forces': 1, 'step': 3}
STEP_KERNELS = ('step_begin', 'forces')


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pair_tally(arr, cell, segments):
    """Algorithmic work of one step's force pass on state `arr`, by the
    per-pair operation tally (10 / 80 / 126 ops per candidate pair that is
    rejected / social / social+contact; 22 / 52 per agent-wall pair without /
    with contact), counting each unordered pair once as the reference does.
    The candidate pairs are those of agents in the same or adjacent cells of
    the binning (origin at the minimum, cell index = floor((p - origin) /
    cell)): a k-d tree lists the pairs closer than 2 sqrt(2) cells, which
    include every such pair, and the cell rule filters them."""
    from scipy.spatial import cKDTree
    p = arr['position']
    l = np.floor((p - p.min(0)) / cell).astype(np.int64)
    pairs = cKDTree(p).query_pairs(r=2.0 * np.sqrt(2.0) * cell * (1.0 + 1e-9),
                                   output_type='ndarray').reshape(-1, 2)
    pairs = pairs[np.all(np.abs(l[pairs[:, 0]] - l[pairs[:, 1]]) <= 1, axis=1)]
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


def step_begin_bytes(n, cells, real_bytes=8):
    """First kernel of the step, per agent.  Binning: read the position 16,
    write the cell index, the slot and the id in the slot list (4 + 4 + 4),
    read-modify-write the cell's count 8.  Own terms: read mass,
    fluctuation sigma, tau_adj, target speed and target 40, the four
    potential values of the bilinear navigation lookup 32, force, velocity
    and target direction (2 reals each); write force and target direction
    (2 reals each).  For the force kernel: write the snapshot (position 16,
    velocity 2 reals) and the slot's neighbour record (position 16, velocity
    2 reals, radius 8, read 8).  156 + 14 reals = 268 B/agent in fp64,
    212 B in fp32; plus clearing the previous step's cell counts, 4 B per
    grid cell."""
    return n * (36 + 40 + 32 + 16 + 32 + 14 * real_bytes) + 4 * cells


def forces_bytes(n, real_bytes=8):
    """Force kernel with the integrator (circular, fused), per agent: read
    position 16, cell index and own slot 4 + 4, radius, mass and the five
    pair and contact parameters 56, target speed 8, velocity, force and
    previous force (2 reals each), and once as a neighbour its record
    (position 16, velocity 2 reals, radius 8); write position 16, velocity,
    previous force and the reset force (2 reals each).  128 + 14 reals =
    240 B/agent in fp64, 184 B in fp32 (neighbour data counted once: every
    agent is read as a neighbour from cache)."""
    return n * (16 + 8 + 56 + 8 + 16 + 8 + 16 + 14 * real_bytes)


def measured_traffic(kernel, threads):
    """HBM bytes per launch of `kernel` (e.g. 'k_forces_direct<double, 0, true>')
    at a grid of `threads` threads, from the committed summary of the two
    rocprofv3 counter runs (FETCH_SIZE and WRITE_SIZE scaled by a 256 MiB
    copy), or None when it is not listed."""
    path = os.path.join(ROOT, 'profiles', 'room20k_hbm_traffic.md')
    try:
        for line in open(path):
            cells = [c.strip() for c in line.strip().strip('|').split('|')]
            if len(cells) == 5 and cells[0] == '`%s`' % kernel and cells[1] == str(threads):
                return int(cells[4])
    except OSError:
        return None
    return None


def host_cpu():
    """Model name and logical CPU count of the host the CPU baseline ran on."""
    model = None
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                model = line.split(':', 1)[1].strip()
                break
    except OSError:
        pass
    return dict(model=model, logical_cpus=os.cpu_count())


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


def bench_ensemble(n_scenarios, steps, loop_steps, stream, tune=True, order_every=10):
    """Batched ensemble of uneven rooms (fp32, fp64 positions): autotuned
    batched steps replayed from a HIP graph versus a per-scenario loop of
    plain GPU steps (one Stepper per scenario, same cell size and
    precision)."""
    import torch
    from crowddynamics_amd import examples as E
    from crowddynamics_amd.autotune import neutral_cell_sizes, max_body_radius, tune_ensemble
    from crowddynamics_amd.ensemble import Ensemble
    from crowddynamics_amd.stepper import Stepper
    log('ensemble: %d scenarios' % n_scenarios)
    scns = E.ensemble(n_scenarios)
    n = sum(s.n for s in scns)
    with torch.cuda.stream(stream):
        sizes = neutral_cell_sizes(max_body_radius(scns))[:3]
        if tune:
            tuned = tune_ensemble(scns, precision='f32', cell_sizes=sizes, steps=5, warmup=2)
        else:
            tuned = dict(cell_size=sizes[0], tile=(0, 0), lanes=4, order_every=order_every,
                         ms_per_step=float('nan'), table=[])
        log('autotune: cell %.3f, tile %s, lanes %d, order_every %d (%.3f ms/step)' % (
            tuned['cell_size'], tuned['tile'], tuned['lanes'], tuned['order_every'],
            tuned['ms_per_step']))
        ens = Ensemble(scns, precision='f32', cell_size=tuned['cell_size'], tile=tuned['tile'],
                       lanes=tuned['lanes'], order_every=tuned['order_every'])
        ens.step(10)
        G = 50
        graph = ens.capture(G)
        reps = max(1, steps // G)
        stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            graph.launch()
        stream.synchronize()
        el = time.perf_counter() - t0
        ens.check()
        kms = ens.timed(20)
        del graph
        batched = n * reps * G / el
        del ens
        log('ensemble batched: %.3e agent-steps/s (%.3f ms/step)' % (batched, 1e3 * el / (reps * G)))
        steppers = [Stepper(s.agents.copy(), segments=s.segments, cell_size=tuned['cell_size'],
                            dt_min=s.dt_min, dt_max=s.dt_max, seed=s.seed, precision='f32',
                            bounds=s.bounds) for s in scns]
        for st in steppers:
            st.step(2)
        stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(loop_steps):
            for st in steppers:
                st.step(1)
        stream.synchronize()
        el_loop = time.perf_counter() - t0
        loop = n * loop_steps / el_loop
        for st in steppers:
            st.check()
        del steppers
    log('ensemble per-scenario loop: %.3e agent-steps/s' % loop)
    return dict(scenarios=n_scenarios, agents=n, precision='f32 (f64 positions and dt)',
                steps=reps * G, value=batched, unit='agent-steps/s',
                ms_per_step=1e3 * el / (reps * G),
                kernel_ms={k: kms[i] / 20 for k, i in KERNELS.items()},
                autotuned=dict(cell_size=tuned['cell_size'],
                               tile=tuned['tile'], lanes=tuned['lanes'],
                               order_every=tuned['order_every'],
                               candidates=[dict(ms=t[0], cell_size=t[1], kernel=t[2])
                                           for t in tuned['table']]),
                per_scenario_loop=dict(value=loop, steps=loop_steps,
                                       ms_per_step=1e3 * el_loop / loop_steps),
                speedup_vs_loop=batched / loop)


def bench_corridor(steps, stream):
    """50k ThreeCircle agents, 300x150 m corridor with 200 obstacles, fp64."""
    import torch
    from crowddynamics_amd import examples as E
    from crowddynamics_amd.stepper import Stepper
    log('corridor: 50k three-circle agents')
    scn = E.corridor()
    G = 50

    def stepper(lanes):
        return Stepper(scn.agents.copy(), segments=scn.segments, cell_size=scn.cell_size,
                       dt_min=scn.dt_min, dt_max=scn.dt_max, seed=scn.seed, precision='f64',
                       bounds=scn.bounds, lanes=lanes)

    with torch.cuda.stream(stream):
        # lanes per agent: 2 or 4, by one graph of G steps each (same results)
        lane_ms = {}
        for lanes in (2, 4):
            st = stepper(lanes)
            st.step(10)
            graph = st.capture(G)
            graph.launch()
            stream.synchronize()
            t0 = time.perf_counter()
            graph.launch()
            stream.synchronize()
            lane_ms[lanes] = 1e3 * (time.perf_counter() - t0) / G
            del graph, st
        lanes = min(lane_ms, key=lane_ms.get)
        log('corridor: ms per step by lanes %s' % lane_ms)
        st = stepper(lanes)
        st.step(10)
        graph = st.capture(G)
        reps = max(1, steps // G)
        stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            graph.launch()
        stream.synchronize()
        el = time.perf_counter() - t0
        st.check()
        kms = st.timed(20)
        del graph
    v = scn.n * reps * G / el
    log('corridor: %.3e agent-steps/s' % v)
    return dict(agents=scn.n, steps=reps * G, value=v, unit='agent-steps/s',
                ms_per_step=1e3 * el / (reps * G), precision='f64', lanes=lanes,
                lane_ms=lane_ms, kernel_ms={k: kms[i] / 20 for k, i in KERNELS.items()})


def bench_adjoint(n, steps, checkpoint):
    """Differentiable rollout: n circular agents, fixed dt 0.005 s, loss =
    mean squared distance to a target trajectory made by the same method
    with tau_0 and kappa scaled by 1.05 and desired speeds by 1 + 0.05 N(0,1);
    gradients for tau_0, mu, kappa, damping and the n desired speeds from
    the adjoint with a checkpoint every `checkpoint` steps."""
    import torch
    from crowddynamics_amd import examples as E
    from crowddynamics_amd.adjoint import Rollout, rollout_positions
    log('adjoint: n=%d, %d steps, checkpoint %d' % (n, steps, checkpoint))
    scn = E.rollout(n=n, steps=steps)
    arr = scn.agents
    dev = 'cuda'
    f64 = torch.float64

    def params(grad=False):
        ps = [torch.tensor(3.0), torch.tensor(1.2e5), torch.tensor(4e4), torch.tensor(500.0),
              torch.as_tensor(np.array(arr['target_velocity'], dtype=np.float64))]
        return [p.to(device=dev, dtype=f64).requires_grad_(grad) for p in ps]

    def rollout(k, lanes):
        return Rollout(arr, steps=k, dt=scn.dt_max, cell_size=scn.cell_size,
                       checkpoint=checkpoint, bounds=scn.bounds, lanes=lanes)

    # lanes per agent: timed on a short rollout (forward + backward)
    lane_ms = {}
    for lanes in (4, 8, 16, 32, 64):
        Rt = rollout(2 * checkpoint, lanes)
        p = params(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        (rollout_positions(Rt, *p) ** 2).sum().backward()
        torch.cuda.synchronize()
        lane_ms[lanes] = 1e3 * (time.perf_counter() - t0) / (2 * checkpoint)
        del Rt
    lanes = min(lane_ms, key=lane_ms.get)
    log('adjoint lanes per agent %d (%s ms/step)' % (lanes, lane_ms))
    R = rollout(steps, lanes)

    rng = np.random.default_rng(0)
    q = params()
    q[0] = q[0] * 1.05
    q[2] = q[2] * 1.05
    q[4] = q[4] * torch.as_tensor(1.0 + 0.05 * rng.standard_normal(n), device=dev)
    with torch.no_grad():
        Xhat = rollout_positions(R, *q)
    # one untimed pass of the whole gradient (torch's loss and reduction
    # kernels load on first use), then the timed one
    p = params(True)
    ((rollout_positions(R, *p)[1:] - Xhat[1:]) ** 2).sum(-1).mean().backward()
    p = params(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    X = rollout_positions(R, *p)
    L = ((X[1:] - Xhat[1:]) ** 2).sum(-1).mean()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    L.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    fwd, bwd = t1 - t0, t2 - t1
    n_par = 4 + n
    log('adjoint: forward %.3f ms/step, backward %.3f ms/step' % (1e3 * fwd / steps,
                                                                   1e3 * bwd / steps))
    return dict(agents=n, steps=steps, checkpoint=checkpoint, dt=scn.dt_max, lanes=lanes,
                loss=float(L.detach()), forward_ms_per_step=1e3 * fwd / steps,
                backward_ms_per_step=1e3 * bwd / steps,
                adjoint_step_ms=1e3 * (fwd + bwd) / steps,
                gradient_s=fwd + bwd,
                grad_norms=dict(tau_0=float(p[0].grad), mu=float(p[1].grad),
                                kappa=float(p[2].grad), damping=float(p[3].grad),
                                v0_l2=float(torch.linalg.vector_norm(p[4].grad))),
                fd_equivalent=dict(
                    note='central differences need 2 forward rollouts per parameter',
                    parameters=n_par, seconds=2 * n_par * fwd,
                    speedup_of_adjoint=2 * n_par * fwd / (fwd + bwd)))


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
    ap.add_argument('--ensemble-scenarios', type=int, default=256)
    ap.add_argument('--ensemble-steps', type=int, default=1000)
    ap.add_argument('--ensemble-loop-steps', type=int, default=20)
    ap.add_argument('--no-ensemble', action='store_true')
    ap.add_argument('--no-corridor', action='store_true')
    ap.add_argument('--no-adjoint', action='store_true')
    ap.add_argument('--adjoint-steps', type=int, default=500)
    ap.add_argument('--lanes', type=int, default=0, choices=(0, 4, 8, 16),
                    help='lanes per agent of the room force kernel (0: autotuned)')
    ap.add_argument('--no-tune', action='store_true',
                    help='skip the autotuners (room: agent-parallel force kernel with 8 '
                         'lanes per agent; ensemble: 4 lanes, cell order, and the smallest '
                         'physics-neutral cell; results are identical for every choice)')
    ap.add_argument('--ensemble-order-every', type=int, default=10,
                    help='with --no-tune: the ensemble force kernel serves agents in cell order '
                         'rebuilt every N steps (0: index order)')
    ap.add_argument('--calibrate-copy', action='store_true',
                    help='first run a plain 256 MiB device copy five times (a kernel of '
                         'known traffic for scaling hardware byte counters)')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU: the per-step update runs on a single MI355X; '
              'sharding a scenario over several GPUs is not built.', flush=True)
        sys.exit(2)

    import torch
    from crowddynamics_amd import _lib
    from crowddynamics_amd import examples as E
    from crowddynamics_amd.stepper import Stepper

    if args.calibrate_copy:
        src = torch.ones(64 << 20, dtype=torch.float32, device='cuda')
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        log('calibration copy: 5 x %d bytes read and written' % (src.numel() * 4))
        del src, dst

    log('building workload: room n=%d' % args.n)
    scn = E.room(n=args.n)
    results = {}
    kernel_ms = {}
    kernel_b2b = {}
    stream = torch.cuda.Stream()
    from crowddynamics_amd.autotune import tune_kernel

    def make(prec, tile, lanes):
        return Stepper(scn.agents.copy(), segments=scn.segments, nav=scn.nav,
                       cell_size=scn.cell_size, dt_min=scn.dt_min, dt_max=scn.dt_max,
                       seed=scn.seed, precision=prec, bounds=scn.bounds, tile=tile, lanes=lanes)

    kern, kern_table = (0, 0, args.lanes or 8), []  # agent-parallel kernel
    if not args.no_tune and not args.lanes:  # lanes per agent, autotuned
        with torch.cuda.stream(stream):
            kern, kern_table = tune_kernel(lambda t, l: make(args.precision, t, l),
                                           kernels=((0, 0, 4), (0, 0, 8), (0, 0, 16)))
    log('force kernel %s (candidates %s)' % (kern, [(t[1], round(t[0], 4)) for t in kern_table]))
    for prec in ('f64', 'f32'):
        with torch.cuda.stream(stream):
            st = make(prec, kern[:2], kern[2])
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
            # per-kernel GPU time, HIP events on the launch stream: inside the
            # step sequence, and over 20 back-to-back launches of each kernel
            # (the step's state is discarded afterwards)
            nt = 50
            ms = st.timed(nt)
            kernel_ms[prec] = {k: ms[i] / nt for k, i in KERNELS.items()}
            b2b = {}
            for k, name in (('forces', 'forces'), ('bin', 'step_begin')):
                st.prepare()
                b2b[name] = st.kernel_timed(k, 20)
            kernel_b2b[prec] = b2b
            results[prec] = dict(value=scn.n * args.steps / el, ms_per_step=1e3 * el / args.steps,
                                 time_tot=state.time_tot, cells=int(state.shape[0] * state.shape[1]),
                                 out=out)
            del graph
        log('%s: %.3e agent-steps/s, %.4f ms/step, kernels %s' % (
            prec, results[prec]['value'], results[prec]['ms_per_step'],
            {k: round(v * 1e3, 2) for k, v in kernel_ms[prec].items()}))

    main_p = args.precision
    km = kernel_b2b[main_p]
    cells = results[main_p]['cells']
    t0_tally = pair_tally(scn.agents, scn.cell_size, scn.segments)
    t1_tally = pair_tally(results[main_p]['out'], scn.cell_size, scn.segments)
    flops_step = 0.5 * (t0_tally['flops'] + t1_tally['flops'])
    cand_step = 0.5 * (t0_tally['candidate_pairs'] + t1_tally['candidate_pairs'])
    rb = 8 if main_p == 'f64' else 4
    per_kernel = {}
    tname = 'double' if main_p == 'f64' else 'float'
    knames = {'step_begin': 'k_step_begin<%s, 0>' % tname,
              'forces': 'k_forces_direct<%s, 0, true>' % tname}
    # grid threads of each launch (the traffic summary is keyed by them)
    f_threads = min(-(-scn.n // 256), 512) * 256  # blocks_f of one scenario
    grid = {'step_begin': f_threads, 'forces': -(-scn.n // (256 // kern[2])) * 256}
    sb = step_begin_bytes(scn.n, cells, rb)
    per_kernel['step_begin'] = dict(bound='hbm', unit='GB/s', ms=km['step_begin'], bytes=sb,
                                    achieved=sb / (km['step_begin'] * 1e-3) / 1e9,
                                    peak=HBM_PEAK_GBS)
    # the force kernel's arithmetic intensity (pair tally flops / bytes) is far
    # below the ridge point (78.6 TFLOP/s / 8 TB/s): memory-side roofline
    fb = forces_bytes(scn.n, rb)
    per_kernel['forces'] = dict(bound='hbm', unit='GB/s', ms=km['forces'], bytes=fb,
                                achieved=fb / (km['forces'] * 1e-3) / 1e9, peak=HBM_PEAK_GBS,
                                flops=flops_step,
                                flops_achieved_tflops=flops_step / (km['forces'] * 1e-3) / 1e12,
                                flops_frac=flops_step / (km['forces'] * 1e-3) / 1e12 /
                                FP64_PEAK_TFLOPS,
                                candidate_pairs_per_s=cand_step / (km['forces'] * 1e-3))
    for k, v in per_kernel.items():
        v['kernel'] = knames[k]
        v['grid_threads'] = grid[k]
        v['frac'] = v['achieved'] / v['peak']
        v['traffic'] = measured_traffic(knames[k], grid[k])
    dom = max(STEP_KERNELS, key=lambda k: km[k])
    rk = per_kernel[dom]
    roofline = dict(kernel=knames[dom], bound=rk['bound'], achieved=rk['achieved'],
                    peak=rk['peak'], unit=rk['unit'], frac=rk['frac'], traffic=rk['traffic'],
                    bytes_per_launch=rk['bytes'], ms_per_launch=rk['ms'])

    ensemble = None
    if not args.no_ensemble:
        ensemble = bench_ensemble(args.ensemble_scenarios, args.ensemble_steps,
                                  args.ensemble_loop_steps, stream, tune=not args.no_tune,
                                  order_every=args.ensemble_order_every)
    corridor = None
    if not args.no_corridor:
        corridor = bench_corridor(300, stream)

    adjoint = None
    if not args.no_adjoint:
        with torch.cuda.stream(stream):
            adjoint = bench_adjoint(4000, args.adjoint_steps, 25)

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
                   all_core=dict(value=vn, cores=nth, steps=sn),
                   host=host_cpu())

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
                   'graph_steps': args.graph_steps, 'force_kernel': list(kern)},
        'roofline': roofline,
        'cpu_baseline': cpu,
        'fp32_value': results['f32']['value'] if main_p == 'f64' else results['f64']['value'],
        'kernel_ms': kernel_b2b,
        'kernel_ms_in_step': kernel_ms,
        'rooflines': per_kernel,
        'pair_tally': t0_tally,
        'ensemble': ensemble,
        'corridor_three_circle': corridor,
        'adjoint': adjoint,
    }
    if cpu:
        line['speedup_vs_cpu_1core'] = r['value'] / cpu['value']
        line['speedup_vs_cpu_allcore'] = r['value'] / cpu['all_core']['value']
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
