"""Python front end of the CPU oracle (oracle/cd_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the reference the GPU results are compared
with or as the timed host baseline.  The product package never imports it.

Functions take the reference's structured agent records (field names of
crowddynamics/simulation/agents.py) and work on contiguous per-field copies,
writing the mutated fields back.

Parity pinning of each oracle part (what checks it besides the GPU):
  wrap_to_pi               Python float '%' semantics, reference KAT for +-pi
                           (tests/test_oracle_math.py)
  distance_circles/_line   brute-force segment sampling, h >= -r_tot invariant,
                           degenerate cases d=0 and sign(0)=0
  distance_three_circles   brute-force min over the 9 part pairs; the x0[j_min]
                           quirk by a constructed case
  power law (circular)     tau plugged back into |x + tau v| = r_tot, tau from
                           numpy polynomial roots, force = -m grad E by central
                           differences of E(tau(x)), reference KATs (co-moving
                           -> 0, head-on -> non-zero), 2 kN truncation
  power law (three-circle) tau = min over the part pairs by polynomial roots
                           where the reference's selection rule reduces to it;
                           the negative-tau lock-in quirk by a constructed case;
                           the shoulder-offset branch (r_ts != 0) by the
                           identity grad = -grad tau(Xp) - 3 r_off / d with
                           grad tau by central differences of numpy roots
                           (the 3 r_off / d term, from the offset's sign and
                           its factor 2, is reference-defined), rotation
                           equivariance and phi -> phi + pi relabelling
                           invariance (tests/test_oracle_pins.py)
  contact                  physical special cases (pure compression mu*delta,
                           pure sliding kappa*delta*v_t, pure approach c*v_n)
                           and Newton's third law with equal parameters
  binning                  numpy stable argsort by cell; pair set equals the
                           brute-force set {i<j : |dlx|<=1 and |dly|<=1}
  integrator               closed-form constant-force and constant-torque
                           trajectories of the reference's Verlet ordering; dt
                           in [dt_min, dt_max]; the reference's validation runs
                           (TestMovement >= 10 m, TestAgentInteraction >= 8 m)
  force_adjust             one agent from rest against the closed form of the
                           Verlet recurrence w_{k+1} = w_k - dt/(2 tau)(w_k +
                           w_{k-1}) by its characteristic roots (position,
                           velocity, previous force to 1e-12)
  torque_adjust            the rotational recurrence solved by eigen-
                           decomposition, across +-pi (short way) and with
                           omega0 != 0 and omega0 = 0
  orientation              bit-equal to math.atan2 on the axes, quadrants and
                           signed zeros
  navigation               exact lattice positions (nearest node) and exact
                           bilinear potentials (gradient mode)
  fluctuation              Philox4x32-10 against its published known-answer
                           vectors; sample moments of the truncated normal
  distance_map (cd_field.c) the exact distances of plane fronts (the
                           hallway's straight targets), the closed-form
                           distance to the exits of a convex room within the
                           first-order scheme's error, the fixed point of the
                           upwind update checked by plain relaxation in numpy
  direction_map            np.gradient itself
  obstacle_handling        a straight wall with a target front along it,
                           where every map is a plane wave and the blended
                           direction has a closed form
  line_intersect,          crossing and parallel segment pairs of known answer;
  nearest neighbours       the k nearest by brute force over all pairs
  herding_interaction      one neighbour heading away (closed form) and the
                           heading-away cone by construction
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, '_build', 'libcd_oracle.so')

_VEC = ('position', 'velocity', 'target_direction', 'force', 'force_prev',
        'position_ls', 'position_rs')
_SCAL = ('radius', 'r_t', 'r_s', 'r_ts', 'mass', 'inertia_rot', 'target_velocity',
         'target_angular_velocity', 'tau_adj', 'k_soc', 'tau_0', 'mu', 'kappa',
         'damping', 'std_rand_force', 'orientation', 'angular_velocity',
         'target_orientation', 'torque', 'torque_prev', 'tau_rot', 'std_rand_torque')
_MUT = ('position', 'velocity', 'target_direction', 'force', 'force_prev',
        'position_ls', 'position_rs', 'orientation', 'angular_velocity',
        'target_orientation', 'torque', 'torque_prev')

_dp = C.POINTER(C.c_double)
_lp = C.POINTER(C.c_int64)


class _Agents(C.Structure):
    _fields_ = ([('n', C.c_int64), ('model', C.c_int32), ('_pad', C.c_int32)]
                + [(f, _dp) for f in ('position', 'velocity', 'target_direction',
                                      'force', 'force_prev', 'position_ls',
                                      'position_rs')]
                + [(f, _dp) for f in ('radius', 'r_t', 'r_s', 'r_ts', 'mass',
                                      'inertia_rot', 'target_velocity',
                                      'target_angular_velocity', 'tau_adj', 'k_soc',
                                      'tau_0', 'mu', 'kappa', 'damping',
                                      'std_rand_force', 'orientation',
                                      'angular_velocity', 'target_orientation',
                                      'torque', 'torque_prev', 'tau_rot',
                                      'std_rand_torque')]
                + [('target', _lp)])


class StepConfig(C.Structure):
    _fields_ = [('cell_size', C.c_double), ('dt_min', C.c_double), ('dt_max', C.c_double),
                ('use_fluct', C.c_int32), ('use_nav', C.c_int32), ('nav_mode', C.c_int32),
                ('use_adjust', C.c_int32), ('use_agent_agent', C.c_int32),
                ('use_obstacle', C.c_int32), ('use_reset', C.c_int32),
                ('orient_velocity', C.c_int32),
                ('seed', C.c_uint64), ('segments', _dp), ('n_segments', C.c_int64),
                ('nav_maps', _dp), ('nav_targets', C.c_int64), ('nav_rows', C.c_int64),
                ('nav_cols', C.c_int64), ('nav_minx', C.c_double), ('nav_miny', C.c_double),
                ('nav_step', C.c_double)]


def build():
    """Compile the oracle library (make in oracle/)."""
    subprocess.run(['make', '-s', '-C', _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_wrap_to_pi.restype = C.c_double
        L.orc_wrap_to_pi.argtypes = [C.c_double]
        L.orc_adaptive_timestep.restype = C.c_double
        L.orc_velocity_verlet_integrator.restype = C.c_double
        L.orc_step.restype = C.c_double
        L.orc_step_omp.restype = C.c_double
        L.orc_iter_pairs.restype = C.c_int64
        _lib = L
    return _lib


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t)


class AgentView:
    """Contiguous per-field copies of a structured agent array, exposed to C."""

    def __init__(self, arr):
        self.arr = arr
        names = arr.dtype.names
        self.model = 1 if 'position_ls' in names else 0
        n = len(arr)
        self.buf = {}
        for f in _VEC:
            self.buf[f] = (np.ascontiguousarray(arr[f], dtype=np.float64).copy()
                           if f in names else np.zeros((n, 2)))
        for f in _SCAL:
            self.buf[f] = (np.ascontiguousarray(arr[f], dtype=np.float64).copy()
                           if f in names else np.zeros(n))
        self.buf['target'] = np.ascontiguousarray(arr['target'], dtype=np.int64).copy()
        s = _Agents()
        s.n = n
        s.model = self.model
        for f in _VEC + _SCAL:
            setattr(s, f, _ptr(self.buf[f]))
        s.target = _ptr(self.buf['target'], _lp)
        self.s = s

    def ref(self):
        return C.byref(self.s)

    def write_back(self):
        for f in _MUT:
            if f in self.arr.dtype.names:
                self.arr[f] = self.buf[f]


def wrap_to_pi(x):
    return lib().orc_wrap_to_pi(float(x))


def distance_circle_line(x, r, p0, p1):
    h = C.c_double()
    n = np.zeros(2)
    x = np.ascontiguousarray(x, dtype=np.float64)
    p0 = np.ascontiguousarray(p0, dtype=np.float64)
    p1 = np.ascontiguousarray(p1, dtype=np.float64)
    lib().orc_distance_circle_line(_ptr(x), C.c_double(r), _ptr(p0), _ptr(p1),
                                   C.byref(h), _ptr(n))
    return h.value, n


def distance_three_circles(x0, r0, x1, r1):
    """(h, normal, r_moment0, r_moment1) for part centres x0, x1 [3, 2]."""
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    x1 = np.ascontiguousarray(x1, dtype=np.float64)
    r0 = np.ascontiguousarray(r0, dtype=np.float64)
    r1 = np.ascontiguousarray(r1, dtype=np.float64)
    h = C.c_double()
    n, m0, m1 = np.zeros(2), np.zeros(2), np.zeros(2)
    lib().orc_distance_three_circles(_ptr(x0), _ptr(r0), _ptr(x1), _ptr(r1), C.byref(h),
                                     _ptr(n), _ptr(m0), _ptr(m1))
    return h.value, n, m0, m1


def pair(arr, i, j):
    """Pair contributions (fi, fj[, mi, mj]) of interaction(i, j) exactly as
    the reference computes them for one pair, or None when h >= 3."""
    v = AgentView(arr)
    fi, fj = np.zeros(2), np.zeros(2)
    if v.model == 0:
        ok = lib().orc_pair_circular(v.ref(), C.c_int64(i), C.c_int64(j), _ptr(fi), _ptr(fj))
        return (fi, fj) if ok else None
    mi, mj = C.c_double(), C.c_double()
    ok = lib().orc_pair_three_circle(v.ref(), C.c_int64(i), C.c_int64(j), _ptr(fi),
                                     _ptr(fj), C.byref(mi), C.byref(mj))
    return (fi, fj, mi.value, mj.value) if ok else None


def add_to_cells(pos, cell_size, capacity=None):
    """Binning contract of cd_oracle.c.  Returns (points_indices, cells_count,
    cells_offset, grid_shape, origin, cell_of)."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n = len(pos)
    if capacity is None:
        span = pos.max(0) - pos.min(0) if n else np.zeros(2)
        capacity = int((span[0] / cell_size + 2) * (span[1] / cell_size + 2)) + 1
    origin = np.zeros(2)
    shape = np.zeros(2, dtype=np.int64)
    cell_of = np.zeros(n, dtype=np.int64)
    pidx = np.zeros(n, dtype=np.int64)
    cnt = np.zeros(capacity, dtype=np.int64)
    off = np.zeros(capacity, dtype=np.int64)
    rc = lib().orc_add_to_cells(_ptr(pos), C.c_int64(n), C.c_double(cell_size),
                                _ptr(origin), _ptr(shape, _lp), _ptr(cell_of, _lp),
                                _ptr(pidx, _lp), _ptr(cnt, _lp), _ptr(off, _lp),
                                C.c_int64(capacity))
    if rc != 0:
        raise ValueError('add_to_cells failed with %d' % rc)
    C_ = int(shape[0] * shape[1])
    return pidx, cnt[:C_].copy(), off[:C_].copy(), (int(shape[0]), int(shape[1])), origin, cell_of


def iter_pairs(pos, cell_size):
    """All (min, max) pairs in the contract's iteration order, [P, 2] int64."""
    pidx, cnt, off, shape, _, _ = add_to_cells(pos, cell_size)
    L = lib()
    total = L.orc_iter_pairs(C.c_int64(shape[0]), C.c_int64(shape[1]), _ptr(pidx, _lp),
                             _ptr(cnt, _lp), _ptr(off, _lp), _lp(), C.c_int64(0))
    pairs = np.zeros((max(total, 1), 2), dtype=np.int64)
    L.orc_iter_pairs(C.c_int64(shape[0]), C.c_int64(shape[1]), _ptr(pidx, _lp),
                     _ptr(cnt, _lp), _ptr(off, _lp), _ptr(pairs, _lp), C.c_int64(total))
    return pairs[:total]


def _call_mut(fname, arr, *args):
    v = AgentView(arr)
    r = getattr(lib(), fname)(v.ref(), *args)
    v.write_back()
    return r


def agent_agent_block_list(arr, cell_size):
    rc = _call_mut('orc_agent_agent_block_list', arr, C.c_double(cell_size))
    if rc != 0:
        raise ValueError('agent_agent_block_list failed with %d' % rc)


def agent_agent_scale(arr, cell_size):
    """agent_agent_block_list, returning per-agent (force_scale, torque_scale):
    the sums of |pair contribution| that bound summation-order rounding."""
    n = len(arr)
    fs = np.zeros(n)
    ts = np.zeros(n)
    rc = _call_mut('orc_agent_agent_scale', arr, C.c_double(cell_size), _ptr(fs), _ptr(ts))
    if rc != 0:
        raise ValueError('agent_agent_block_list failed with %d' % rc)
    return fs, ts


def agent_obstacle(arr, segments):
    seg = np.ascontiguousarray(_segments_array(segments))
    _call_mut('orc_agent_obstacle', arr, _ptr(seg), C.c_int64(len(seg)))


def _segments_array(segments):
    s = np.asarray(segments)
    if s.dtype.names:
        out = np.zeros((len(s), 4))
        out[:, 0:2] = s['p0']
        out[:, 2:4] = s['p1']
        return out
    return np.asarray(s, dtype=np.float64).reshape(-1, 4)


def force_adjust_agents(arr):
    _call_mut('orc_force_adjust_agents', arr)


def torque_adjust_agents(arr):
    _call_mut('orc_torque_adjust_agents', arr)


def orient_towards_target_direction(arr):
    _call_mut('orc_orient_towards_target_direction', arr)


def adaptive_timestep(arr, dt_min, dt_max):
    v = AgentView(arr)
    return lib().orc_adaptive_timestep(v.ref(), C.c_double(dt_min), C.c_double(dt_max))


def velocity_verlet_integrator(arr, dt_min, dt_max):
    return _call_mut('orc_velocity_verlet_integrator', arr, C.c_double(dt_min),
                     C.c_double(dt_max))


def reset(arr):
    _call_mut('orc_reset', arr)


def navigation(arr, maps, minx, miny, step, mode=0):
    """Navigation update.  mode 0: maps [T, rows, cols, 2] direction maps read at
    the nearest lower node (reference).  mode 1: maps [T, rows, cols] potentials
    read by bilinear gradient."""
    maps = np.ascontiguousarray(maps, dtype=np.float64)
    T, rows, cols = maps.shape[:3]
    fn = 'orc_navigation_nearest' if mode == 0 else 'orc_navigation_bilinear'
    _call_mut(fn, arr, _ptr(maps), C.c_int64(T), C.c_int64(rows), C.c_int64(cols),
              C.c_double(minx), C.c_double(miny), C.c_double(step))


def fluctuation(arr, seed, step_index):
    _call_mut('orc_fluctuation', arr, C.c_uint64(seed), C.c_uint64(step_index))


def fluctuation_noise(seed, step_index, i, mass, scale_f, inertia=0.0, scale_t=0.0):
    f = np.zeros(2)
    m = C.c_double()
    lib().orc_fluctuation_noise(C.c_uint64(seed), C.c_uint64(step_index), C.c_int64(i),
                                C.c_double(mass), C.c_double(scale_f), C.c_double(inertia),
                                C.c_double(scale_t), _ptr(f), C.byref(m))
    return f, m.value


def philox4x32(ctr, key):
    c = (C.c_uint32 * 4)(*[int(x) & 0xffffffff for x in ctr])
    k = (C.c_uint32 * 2)(*[int(x) & 0xffffffff for x in key])
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32(c, k, o)
    return [int(x) for x in o]


class Scenario:
    """Static step inputs: walls, navigation field and tree switches."""

    def __init__(self, cell_size=3.6, dt_min=0.01, dt_max=0.01, seed=0,
                 segments=None, nav_maps=None, nav_mode=0, nav_origin=(0.0, 0.0),
                 nav_step=0.1, fluct=True, adjust=True, agent_agent=True,
                 obstacle=True, reset=True, navigation=True, orient_velocity=False):
        self.segments = (np.ascontiguousarray(_segments_array(segments))
                         if segments is not None and len(segments) else np.zeros((0, 4)))
        self.nav_maps = (np.ascontiguousarray(nav_maps, dtype=np.float64)
                         if nav_maps is not None else None)
        cfg = StepConfig()
        cfg.cell_size, cfg.dt_min, cfg.dt_max = cell_size, dt_min, dt_max
        cfg.use_fluct = int(fluct)
        cfg.use_nav = int(navigation and self.nav_maps is not None)
        cfg.nav_mode = int(nav_mode)
        cfg.use_adjust = int(adjust)
        cfg.use_agent_agent = int(agent_agent)
        cfg.use_obstacle = int(obstacle)
        cfg.use_reset = int(reset)
        cfg.orient_velocity = int(orient_velocity)
        cfg.seed = seed
        cfg.segments = _ptr(self.segments)
        cfg.n_segments = len(self.segments)
        if self.nav_maps is not None:
            cfg.nav_maps = _ptr(self.nav_maps)
            cfg.nav_targets, cfg.nav_rows, cfg.nav_cols = self.nav_maps.shape[:3]
        cfg.nav_minx, cfg.nav_miny = nav_origin
        cfg.nav_step = nav_step
        self.cfg = cfg


def run(arr, scenario, steps, first_step=0, threads=None):
    """Advance `arr` in place by `steps` steps; returns the dt of each step.
    threads=None runs the single-thread oracle; an int runs the OpenMP
    owner-computes variant (circular only) on that many threads."""
    v = AgentView(arr)
    dts = np.zeros(max(steps, 1))
    if threads is None:
        rc = lib().orc_run(v.ref(), C.byref(scenario.cfg), C.c_uint64(first_step),
                           C.c_int64(steps), _ptr(dts))
    else:
        rc = lib().orc_run_omp(v.ref(), C.byref(scenario.cfg), C.c_uint64(first_step),
                               C.c_int64(steps), C.c_int(threads), _ptr(dts))
    v.write_back()
    if rc != 0:
        raise ValueError('oracle run failed with %d' % rc)
    return dts[:steps]


def max_threads():
    return lib().orc_max_threads()


# ---------------------------------------------------------------- fields --

def distance_map(phi, dx, mask=None):
    """scikit-fmm's distance(phi, dx) restated (cd_field.c): signed distance
    to the zero contour of phi [rows, cols]; NaN at masked and unreachable
    nodes."""
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    rows, cols = phi.shape
    out = np.empty_like(phi)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    rc = lib().orc_distance_map(_ptr(phi), None if m is None else m.ctypes.data_as(C.c_void_p),
                                C.c_int64(rows), C.c_int64(cols), C.c_double(dx), _ptr(out))
    if rc != 0:
        raise ValueError('distance_map: no node borders the contour')
    return out


def direction_map(dmap):
    """(ex, ey): the normalised np.gradient of dmap (cd_field.c)."""
    d = np.ascontiguousarray(dmap, dtype=np.float64)
    ex = np.empty_like(d)
    ey = np.empty_like(d)
    lib().orc_direction_map(_ptr(d), C.c_int64(d.shape[0]), C.c_int64(d.shape[1]), _ptr(ex),
                            _ptr(ey))
    return ex, ey


def obstacle_handling(dobs, ox, oy, tx, ty, radius, strength):
    """Blended and normalised directions (cd_field.c), new arrays."""
    tx = np.array(tx, dtype=np.float64, order='C')
    ty = np.array(ty, dtype=np.float64, order='C')
    args = [np.ascontiguousarray(a, dtype=np.float64) for a in (dobs, ox, oy)]
    lib().orc_obstacle_handling(*[_ptr(a) for a in args], _ptr(tx), _ptr(ty),
                                C.c_int64(tx.size), C.c_double(radius), C.c_double(strength))
    return tx, ty


# ------------------------------------------------------- collective motion --

def line_intersect(x0, x1, y0, y1):
    """line_intersect, crowddynamics/core/geom2D.py:38-59."""
    ux, uy = x1[0] - x0[0], x1[1] - x0[1]
    vx, vy = y1[0] - y0[0], y1[1] - y0[1]
    bx, by = y0[0] - x0[0], y0[1] - x0[1]
    d = ux * vy - uy * vx
    if d == 0:
        return False
    t0 = bx * vy - by * vx
    t1 = bx * uy - by * ux
    return 0 <= t0 / d <= 1 and 0 <= t1 / d <= 1


def nearest_neighbors(position, sight, k, segments):
    """The set computed by find_nearest_neighbors
    (core/steering/collective_motion.py:69-107), by brute force over all
    pairs: the k nearest agents strictly within sight whose line of sight
    crosses no segment ([W, 4]), nearest first, ties to the smaller index,
    -1 where missing."""
    n = len(position)
    out = np.full((n, k), -1, dtype=np.int64)
    segs = np.asarray(segments, dtype=np.float64).reshape(-1, 4)
    for i in range(n):
        cand = []
        for j in range(n):
            if j == i:
                continue
            l = float(np.hypot(position[i, 0] - position[j, 0], position[i, 1] - position[j, 1]))
            if not l < sight:
                continue
            if any(line_intersect(position[i], position[j], s[0:2], s[2:4]) for s in segs):
                continue
            cand.append((l, j))
        cand.sort()
        for m, (_, j) in enumerate(cand[:k]):
            out[i, m] = j
    return out


def _normalize(v):
    l = float(np.hypot(v[0], v[1]))
    return v / l if l else v


def herding_interaction(is_herding, position, velocity, neighbors, weight_position, phi):
    """herding_interaction (collective_motion.py:110-154) with
    herding_relationship (:26-58), over the neighbour lists in their order."""
    n = len(position)
    new = np.zeros_like(position)
    has = np.zeros(n, dtype=bool)
    cos_phi = np.cos(phi)
    for i in range(n):
        if not is_herding[i]:
            continue
        mp = np.zeros(2)
        mv = np.zeros(2)
        num = 0
        for j in neighbors[i]:
            if j < 0:
                continue
            v1, v2 = velocity[i], velocity[j]
            if np.hypot(*v1) == 0 or np.hypot(*v2) == 0:
                continue
            e = _normalize(position[j] - position[i])
            n2 = _normalize(v2)
            cj = -(e[0] * n2[0] + e[1] * n2[1])
            if cos_phi < cj < 1.0:
                continue
            mp += position[j]
            mv += velocity[j]
            num += 1
        if num > 0:
            e0 = _normalize(mp / num - position[i])
            e1 = _normalize(mv)
            new[i] = _normalize(weight_position * e0 + (1 - weight_position) * e1)
            has[i] = True
    return new, has
