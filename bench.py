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
    """HBM bytes per launch of `kernel` (e.g. 'k_forces_direct<double, 0, true, true>')
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


def pair_tally_three_circle(arr, cell, segments):
    """pair_tally for three-circle agents, by this build's count of the
    three-circle pair (same convention: each add/sub/mul/div/sqrt/exp/hypot
    is one op): 9 part distances (4 ops each) and the sight test, 38 ops
    when rejected (h >= 3 at the closest part pair); the time to collision
    of 9 part pairs (16 ops each), the gradient, magnitude, shoulder offset
    and truncation, 230 ops when social; 280 with contact and torque.
    Walls: 3 parts x the circle-segment distance, 66 ops, 96 with contact."""
    from scipy.spatial import cKDTree
    p = arr['position']
    l = np.floor((p - p.min(0)) / cell).astype(np.int64)
    pairs = cKDTree(p).query_pairs(r=2.0 * np.sqrt(2.0) * cell * (1.0 + 1e-9),
                                   output_type='ndarray').reshape(-1, 2)
    pairs = pairs[np.all(np.abs(l[pairs[:, 0]] - l[pairs[:, 1]]) <= 1, axis=1)]
    i, j = pairs[:, 0], pairs[:, 1]
    parts = (arr['position'], arr['position_ls'], arr['position_rs'])
    radii = (arr['r_t'], arr['r_s'], arr['r_s'])
    h = np.full(len(pairs), np.inf)
    for a in range(3):
        for b in range(3):
            d = np.hypot(*(parts[a][i] - parts[b][j]).T) - (radii[a][i] + radii[b][j])
            h = np.minimum(h, d)
    n_rej = int(np.sum(h >= 3.0))
    n_con = int(np.sum(h < 0.0))
    n_soc = len(h) - n_rej - n_con
    seg = np.concatenate([segments['p0'], segments['p1']], 1) if len(segments) else np.zeros((0, 4))
    from crowddynamics_amd.examples import _seg_dist
    w_con = 0
    if len(seg):
        dmin = np.full((len(p), len(seg)), np.inf)
        for a in range(3):
            dmin = np.minimum(dmin, _seg_dist(parts[a][:, 0], parts[a][:, 1], seg)
                              - radii[a][:, None])
        w_con = int(np.sum(dmin < 0))
    w_all = len(p) * len(seg)
    flops = 38 * n_rej + 230 * n_soc + 280 * n_con + 66 * (w_all - w_con) + 96 * w_con
    return dict(candidate_pairs=len(h), social_pairs=n_soc, contact_pairs=n_con,
                wall_pairs=w_all, wall_contacts=w_con, flops=flops)


def _window(st, graph, G, reps, stream, restore):
    """Per-kernel GPU time and throughput over the same steps: the window of
    reps * G steps is first run eagerly with HIP events around every kernel
    (cd_step_timed), then the state is restored and the same window is
    replayed from the graph and timed on the host.  Returns (kernel ms per
    step by name, elapsed seconds of the replay)."""
    import torch
    kms = st.timed(reps * G)
    restore()
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        graph.launch()
    stream.synchronize()
    el = time.perf_counter() - t0
    torch.cuda.synchronize()
    return {k: kms[i] / (reps * G) for k, i in KERNELS.items()}, el


def _snapshot(st, array):
    """A restore() for `st`: its agents' records and step counter now."""
    arr = st.download(array.copy())
    step = st.state().step

    def restore():
        st.agents.upload(arr)
        st.prepare(step)
    return restore


def _force_roofline(n, ms, bytes_, flops, cand):
    return dict(bound='hbm', unit='GB/s', ms=ms, bytes=bytes_,
                achieved=bytes_ / (ms * 1e-3) / 1e9, peak=HBM_PEAK_GBS,
                frac=bytes_ / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                flops=flops, flops_achieved_tflops=flops / (ms * 1e-3) / 1e12,
                flops_frac=flops / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                candidate_pairs_per_s=cand / (ms * 1e-3))


def bench_ensemble(n_scenarios, steps, loop_steps, stream, tune=True, order_every=10):
    """Batched ensemble of uneven rooms (fp32, fp64 positions): autotuned
    batched steps replayed from a HIP graph versus a per-scenario loop of
    plain GPU steps (one Stepper per scenario, same cell size and
    precision), also replayed from a HIP graph."""
    import torch
    from crowddynamics_amd import examples as E
    from crowddynamics_amd.autotune import neutral_cell_sizes, max_body_radius, tune_ensemble
    from crowddynamics_amd.ensemble import Ensemble
    from crowddynamics_amd.stepper import Stepper, capture_many
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
        reps = max(1, steps // G)
        arr0 = np.concatenate([s.agents for s in scns])
        restore = _snapshot(ens, arr0)
        tally0 = _ensemble_tally(ens, scns, arr0, tuned['cell_size'])
        graph = ens.capture(G)
        restore()
        kms, el = _window(ens, graph, G, reps, stream, restore)
        ens.check()
        tally1 = _ensemble_tally(ens, scns, arr0, tuned['cell_size'])
        del graph
        batched = n * reps * G / el
        del ens
        log('ensemble batched: %.3e agent-steps/s (%.3f ms/step), kernels %s' % (
            batched, 1e3 * el / (reps * G), {k: round(v, 4) for k, v in kms.items()}))
        steppers = [Stepper(s.agents.copy(), segments=s.segments, cell_size=tuned['cell_size'],
                            dt_min=s.dt_min, dt_max=s.dt_max, seed=s.seed, precision='f32',
                            bounds=s.bounds) for s in scns]
        for st in steppers:
            st.step(2)
        GL = 5
        loop_graph = capture_many(steppers, GL)
        loop_graph.launch()
        stream.synchronize()
        loop_reps = max(1, loop_steps // GL)
        t0 = time.perf_counter()
        for _ in range(loop_reps):
            loop_graph.launch()
        stream.synchronize()
        el_loop = time.perf_counter() - t0
        loop = n * loop_reps * GL / el_loop
        for st in steppers:
            st.check()
        del loop_graph, steppers
    log('ensemble per-scenario loop (graph): %.3e agent-steps/s' % loop)
    flops = 0.5 * (tally0['flops'] + tally1['flops'])
    cand = 0.5 * (tally0['candidate_pairs'] + tally1['candidate_pairs'])
    roof = _force_roofline(n, kms['forces'], forces_bytes(n, 4), flops, cand)
    return dict(scenarios=n_scenarios, agents=n, precision='f32 (f64 positions and dt)',
                steps=reps * G, value=batched, unit='agent-steps/s',
                ms_per_step=1e3 * el / (reps * G),
                kernel_ms=kms,
                kernel_ms_note='HIP events around each kernel over the same %d steps '
                               '(eager), from the same state as the graph replay' % (reps * G),
                roofline_forces=roof,
                pair_tally=dict(first_step=tally0, last_step=tally1),
                autotuned=dict(cell_size=tuned['cell_size'],
                               tile=tuned['tile'], lanes=tuned['lanes'],
                               order_every=tuned['order_every'],
                               candidates=[dict(ms=t[0], cell_size=t[1], kernel=t[2])
                                           for t in tuned['table']]),
                per_scenario_loop=dict(value=loop, steps=loop_reps * GL,
                                       ms_per_step=1e3 * el_loop / (loop_reps * GL),
                                       note='one Stepper per scenario, one graph of %d rounds '
                                            'of a step of every scenario' % GL),
                speedup_vs_loop=batched / loop)


def _ensemble_tally(ens, scns, arr0, cell):
    """Pair tally of the batch's current state, summed over the scenarios."""
    arr = ens.download(arr0.copy())
    tot = {}
    for k, s in enumerate(scns):
        t = pair_tally(arr[ens.offsets[k]:ens.offsets[k + 1]], cell, s.segments)
        for key, v in t.items():
            tot[key] = tot.get(key, 0) + v
    return tot


def bench_corridor(steps, stream):
    """50k ThreeCircle agents, 300x150 m corridor with 200 obstacles, fp64,
    target orientation from the velocity."""
    import torch
    from crowddynamics_amd import _lib
    from crowddynamics_amd import examples as E
    from crowddynamics_amd.stepper import ALL_NODES, Stepper
    log('corridor: 50k three-circle agents')
    scn = E.corridor()
    G = 50
    flags = ALL_NODES | (_lib.CD_USE_ORIENT_VELOCITY if scn.extra.get('orient') == 'velocity'
                         else 0)

    def stepper(lanes, precision='f64'):
        return Stepper(scn.agents.copy(), segments=scn.segments, cell_size=scn.cell_size,
                       dt_min=scn.dt_min, dt_max=scn.dt_max, seed=scn.seed, precision=precision,
                       bounds=scn.bounds, lanes=lanes, flags=flags)

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
        reps = max(1, steps // G)
        restore = _snapshot(st, scn.agents)
        tally0 = pair_tally_three_circle(st.download(scn.agents.copy()), scn.cell_size,
                                         scn.segments)
        graph = st.capture(G)
        restore()
        kms, el = _window(st, graph, G, reps, stream, restore)
        st.check()
        tally1 = pair_tally_three_circle(st.download(scn.agents.copy()), scn.cell_size,
                                         scn.segments)
        del graph, st
        # the same workload in the fp32 mode (float velocity, force and pair
        # arithmetic; fp64 positions, orientation and torque), graph-replayed
        st = stepper(lanes, 'f32')
        st.step(10)
        graph = st.capture(G)
        graph.launch()
        stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            graph.launch()
        stream.synchronize()
        el32 = time.perf_counter() - t0
        st.check()
        del graph, st
    v = scn.n * reps * G / el
    log('corridor: %.3e agent-steps/s, kernels %s' % (v, {k: round(x, 4) for k, x in kms.items()}))
    flops = 0.5 * (tally0['flops'] + tally1['flops'])
    cand = 0.5 * (tally0['candidate_pairs'] + tally1['candidate_pairs'])
    # three-circle force kernel bytes per agent: the circular model's plus the
    # shoulders' snapshot read as a neighbour (2 x 16 + offset 16), r_t, r_s,
    # r_ts, inertia, orientation, angular velocity, torque, previous torque,
    # tau_rot, target orientation and target angular velocity read (11 x 8),
    # orientation, angular velocity, torques and both shoulders written
    # (4 x 8 + 2 x 16)
    fb = forces_bytes(scn.n, 8) + scn.n * (48 + 88 + 64)
    roof = _force_roofline(scn.n, kms['forces'], fb, flops, cand)
    return dict(agents=scn.n, steps=reps * G, value=v, unit='agent-steps/s',
                ms_per_step=1e3 * el / (reps * G), precision='f64', lanes=lanes,
                fp32_value=scn.n * reps * G / el32, fp32_ms_per_step=1e3 * el32 / (reps * G),
                orientation='angle(velocity)' if flags & _lib.CD_USE_ORIENT_VELOCITY
                else 'angle(target_direction)',
                lane_ms=lane_ms, kernel_ms=kms,
                kernel_ms_note='HIP events around each kernel over the same %d steps '
                               '(eager), from the same state as the graph replay' % (reps * G),
                roofline_forces=roof, pair_tally=dict(first_step=tally0, last_step=tally1))


def _adjoint_window(arr, scn, steps, checkpoint, lanes, seed=0):
    """A rollout of `steps` steps of the differentiable workload, its target
    trajectory (tau_0 and kappa x 1.05, desired speeds x (1 + 0.05 N(0, 1)))
    and the parameter constructor."""
    import torch
    from crowddynamics_amd.adjoint import Rollout, rollout_positions
    dev, f64 = 'cuda', torch.float64
    n = len(arr)

    def params(grad=False):
        ps = [torch.tensor(3.0), torch.tensor(1.2e5), torch.tensor(4e4), torch.tensor(500.0),
              torch.as_tensor(np.array(arr['target_velocity'], dtype=np.float64))]
        return [p.to(device=dev, dtype=f64).requires_grad_(grad) for p in ps]

    R = Rollout(arr, steps=steps, dt=scn.dt_max, cell_size=scn.cell_size,
                checkpoint=checkpoint, bounds=scn.bounds, lanes=lanes)
    rng = np.random.default_rng(seed)
    q = params()
    q[0] = q[0] * 1.05
    q[2] = q[2] * 1.05
    q[4] = q[4] * torch.as_tensor(1.0 + 0.05 * rng.standard_normal(n), device=dev)
    with torch.no_grad():
        Xhat = rollout_positions(R, *q)
    return R, Xhat, params


def _loss(X, Xhat):
    return ((X[1:] - Xhat[1:]) ** 2).sum(-1).mean()


def _gradient(R, Xhat, params):
    """Loss and the adjoint gradient (tau_0, mu, kappa, damping as global
    scalars; the per-agent desired speeds)."""
    from crowddynamics_amd.adjoint import rollout_positions
    p = params(True)
    L = _loss(rollout_positions(R, *p), Xhat)
    L.backward()
    return float(L.detach()), [t.grad.detach() for t in p]


def _gradient_summary(g):
    import torch
    return dict(tau_0=float(g[0]), mu=float(g[1]), kappa=float(g[2]), damping=float(g[3]),
                v0_l2_norm=float(torch.linalg.vector_norm(g[4])),
                v0_max_abs=float(g[4].abs().max()))


def adjoint_checks(arr, scn, lanes, windows=(5, 10, 25, 50, 100),
                   hs=(1e-4, 1e-5, 1e-6, 1e-7, 1e-8),
                   checkpoint=25):
    """On the differentiable workload, for rollouts of each length in
    `windows`: the dot-product test of the adjoint against the forward-mode
    tangent (both exact derivatives of the same rollout: they agree to
    rounding whatever the gradient's size), central differences of the loss
    in tau_0 at the steps `hs`, the smallest agent-agent gap over the window
    (negative: contacts), and the adjoint gradient."""
    import torch
    from crowddynamics_amd.adjoint import rollout_positions
    out = []
    for W in windows:
        R, Xhat, params = _adjoint_window(arr, scn, W, min(checkpoint, W), lanes)
        L, g = _gradient(R, Xhat, params)
        # dot-product test along a random tangent of all parameters and of
        # the initial position and velocity
        n = len(arr)
        R.set_parameters(*params())
        X = R.forward().clone().requires_grad_(True)
        _loss(X, Xhat).backward()
        gX = X.grad.detach()
        gp, lx0, lv0 = R.backward(gX)
        rng = np.random.default_rng(W)
        dp = torch.as_tensor(rng.standard_normal((n, 5)) *
                             (1e-3 * np.array([3.0, 1.2e5, 4e4, 500.0, 1.15])), device='cuda')
        dx0 = torch.as_tensor(rng.standard_normal((n, 2)) * 1e-3, device='cuda')
        dv0 = torch.as_tensor(rng.standard_normal((n, 2)) * 1e-3, device='cuda')
        dX = R.tangent(dp, dx0, dv0)
        terms = gX * dX
        lhs = float(terms.sum())
        rhs = float((gp * dp).sum() + (lx0 * dx0).sum() + (lv0 * dv0).sum())
        dot_rel = abs(lhs - rhs) / max(abs(lhs), float(terms.abs().sum()), 1e-300)
        # central differences in tau_0
        fd = {}
        for h in hs:
            vals = []
            for sgn in (1.0, -1.0):
                p = params()
                p[0] = p[0] + sgn * h
                with torch.no_grad():
                    vals.append(float(_loss(rollout_positions(R, *p), Xhat)))
            fd['%g' % h] = (vals[0] - vals[1]) / (2 * h)
        with torch.no_grad():
            Xc = rollout_positions(R, *params())
            rr = torch.as_tensor(np.array(arr['radius']), device='cuda')
            gap = float('inf')
            for k in range(0, W + 1, max(1, W // 10)):
                d = torch.cdist(Xc[k], Xc[k]) - (rr[:, None] + rr[None, :])
                d.fill_diagonal_(float('inf'))
                gap = min(gap, float(d.min()))
        adj = float(g[0])

        def rel_diff(h):
            floor = 100 * np.finfo(np.float64).eps * abs(L) / h
            return abs(fd['%g' % h] - adj) / (abs(adj) + floor + 1e-300)
        agree = rel_diff(1e-6) <= 1e-4
        best_h = min(hs, key=rel_diff)
        rec = dict(steps=W, loss=L, dot_product=dict(tangent=lhs, adjoint=rhs,
                                                      rel_diff=dot_rel),
                   tau_0_adjoint=adj, tau_0_central_fd=fd,
                   fd_rel_diff={'%g' % h: rel_diff(h) for h in hs},
                   fd_agrees_at_h_1e_6=bool(agree),
                   fd_best=dict(h=best_h, rel_diff=rel_diff(best_h),
                                converged=bool(rel_diff(best_h) <= 1e-4)),
                   min_gap_m=gap, gradient=_gradient_summary(g))
        log('adjoint window %d steps: dot-product rel diff %.2e, dL/dtau_0 adjoint %.6e, '
            'fd(1e-6) %.6e (%s), min gap %.3f m' % (W, dot_rel, adj, fd['1e-06'],
                                                  'agrees' if agree else 'differs', gap))
        out.append(rec)
        del R
    return out


def bench_adjoint(n, steps, checkpoint):
    """Differentiable rollout: n circular agents, fixed dt 0.005 s, loss =
    mean squared distance to a target trajectory made by the same method
    with tau_0 and kappa scaled by 1.05 and desired speeds by 1 + 0.05 N(0,1);
    gradients for tau_0, mu, kappa, damping and the n desired speeds from
    the adjoint with a checkpoint every `checkpoint` steps.  Reports the
    timing of the whole `steps`-step gradient, the dot-product and
    finite-difference checks on windows of 10 to 100 steps, and the gradient
    of the longest window whose central difference (h = 1e-6) agrees."""
    import torch
    from crowddynamics_amd import examples as E
    from crowddynamics_amd.adjoint import rollout_positions
    log('adjoint: n=%d, %d steps, checkpoint %d' % (n, steps, checkpoint))
    scn = E.rollout(n=n, steps=steps)
    arr = scn.agents

    # lanes per agent: timed on a short rollout (forward + backward)
    lane_ms = {}
    for lanes in (4, 8, 16, 32, 64):
        Rt, Xt, pt = _adjoint_window(arr, scn, 2 * checkpoint, checkpoint, lanes)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _gradient(Rt, Xt, pt)
        torch.cuda.synchronize()
        lane_ms[lanes] = 1e3 * (time.perf_counter() - t0) / (2 * checkpoint)
        del Rt
    lanes = min(lane_ms, key=lane_ms.get)
    log('adjoint lanes per agent %d (%s ms/step)' % (lanes, lane_ms))
    R, Xhat, params = _adjoint_window(arr, scn, steps, checkpoint, lanes)
    # one untimed pass of the whole gradient (torch's loss and reduction
    # kernels load on first use), then the timed one
    _gradient(R, Xhat, params)
    p = params(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    X = rollout_positions(R, *p)
    L = _loss(X, Xhat)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    L.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    fwd, bwd = t1 - t0, t2 - t1
    # one central difference (two forward rollouts of the same length), timed
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    for sgn in (1.0, -1.0):
        q = params()
        q[0] = q[0] + sgn * 1e-6
        with torch.no_grad():
            float(_loss(rollout_positions(R, *q), Xhat))
    torch.cuda.synchronize()
    fd_one = time.perf_counter() - t3
    del R
    log('adjoint: forward %.3f ms/step, backward %.3f ms/step' % (1e3 * fwd / steps,
                                                                   1e3 * bwd / steps))
    checks = adjoint_checks(arr, scn, lanes)
    good = [c for c in checks if c['fd_agrees_at_h_1e_6']]
    best = max(good, key=lambda c: c['steps']) if good else None
    conv = [c for c in checks if c['fd_best']['converged']]
    best_conv = max(conv, key=lambda c: c['steps']) if conv else None
    return dict(agents=n, steps=steps, checkpoint=checkpoint, dt=scn.dt_max, lanes=lanes,
                lane_ms=lane_ms, loss=float(L.detach()),
                forward_ms_per_step=1e3 * fwd / steps,
                backward_ms_per_step=1e3 * bwd / steps,
                adjoint_step_ms=1e3 * (fwd + bwd) / steps,
                gradient_s=fwd + bwd,
                one_central_difference_s=fd_one,
                gradient_full_window=dict(
                    steps=steps, values=_gradient_summary([t.grad for t in p]),
                    note='not confirmed by finite differences at this horizon: the '
                         'time-to-collision force grows without bound near grazing '
                         'pairs, so the loss is not smooth at h = 1e-6 over long windows'),
                checks=checks,
                fd_confirmed_window=(dict(steps=best['steps'], gradient=best['gradient'],
                                          loss=best['loss'],
                                          tau_0_central_fd=best['tau_0_central_fd']['1e-06'])
                                     if best else None),
                fd_converged_window=(dict(steps=best_conv['steps'],
                                          h=best_conv['fd_best']['h'],
                                          rel_diff=best_conv['fd_best']['rel_diff'],
                                          gradient=best_conv['gradient'])
                                     if best_conv else None),
                dot_product_max_rel_diff=max(c['dot_product']['rel_diff'] for c in checks))


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
    # the one-scenario instances (last template argument) of the fused step's kernels
    knames = {'step_begin': 'k_step_begin<%s, 0, true, true>' % tname,
              'forces': 'k_forces_direct<%s, 0, true, true>' % tname}
    # grid threads of each launch (the traffic summary is keyed by them): the
    # first kernel of one scenario with fluctuation or navigation runs
    # 4 x blocks_f blocks of 128 threads (binning and own-terms waves)
    blocks_f = min(-(-scn.n // 256), 512)
    grid = {'step_begin': 4 * blocks_f * 128, 'forces': -(-scn.n // (256 // kern[2])) * 256}
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
