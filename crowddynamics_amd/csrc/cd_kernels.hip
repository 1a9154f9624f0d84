// MI355X (gfx950) kernels of the crowddynamics per-timestep agent update and
// the C ABI declared in include/cdamd.h.
//
// Workspace layout, binning contract, reductions and the binning kernels
// K_a / K_b are in cd_common.h.
//
// One step is two kernels, each launched once over all scenarios of a batch
// (a single simulation is a batch of one):
//   K_a  k_step_begin     per scenario: the reduction of the state the step
//                         starts from (bounding box, max |v|, max v0), the
//                         grid, dt; per agent: a snapshot of its state, the
//                         step's own terms (fluctuation, navigation,
//                         orientation, adjusting) and its insert into its
//                         cell's slot list (a single scenario: binning and
//                         own terms on two waves per 64 agents)
//   K_e  k_forces_direct  L = 2 to 64 lanes per agent: the pairs of its
//        (k_forces_tiled) 3 x 3 cell stencil and the walls, summed exactly;
//                         then Verlet and Reset of the agent, the block's
//                         reduction of the new state merged into the
//                         scenario's slot for the next step, the step counted
// Every block serves a single scenario (block maps built by
// cd_batch_prepare).  Pair forces are owner-computes: each agent sums its
// own side of every pair, so forces need no atomics, and the sums are exact,
// so results do not depend on the lanes, tiles or insertion order.  K_e
// reads neighbours from K_a's snapshot, since it moves agents as it goes.
// K_b (k_cell_gather, cd_common.h) serves the adjoint only.

#include "cd_common.h"

namespace {

// ---------------------------------------------------------------------------
// Block-list arrays of cd_add_to_cells (single scenario; counting sort with
// offsets)
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_bl_count(const double2 *pos, double cell, Work W,
                                                  Batch B) {
  const int2 sk = block_f(B);
  const ScnDesc d = desc_of(B, sk.x);
  ScnState *ss = &B.sst[sk.x];
  const Red r = reduce_partials(W, d);
  const Grid g = grid_of(r, cell, d.ccap);
  if (sk.y == 0 && threadIdx.x == 0) {
    publish_grid(W, ss, r, g, true, d.a1 > d.a0, 0);
    W.st->scan_epoch = (W.st->scan_epoch + 1u) & 0x3fffffffu;
    if (W.st->scan_epoch == 0u) W.st->scan_epoch = 1u;
    W.st->scan_ticket = 0u;
  }
  if (!g.ok) return;
  const int stride = d.nf * blockDim.x;
  for (int i = d.a0 + sk.y * blockDim.x + threadIdx.x; i < d.a1; i += stride) {
    const int c = cell_of_pos(pos[i], g, cell);
    if (c < 0) {
      atomicOr(&ss->oob[0], 1);
      continue;
    }
    W.cell_of[i] = c;
    W.rank[i] = atomicAdd(&W.count[c], 1);
  }
}

// Exclusive scan of the counts, single pass with decoupled look-back.  Each
// workgroup takes the next tile of kScanTile cells (dynamic ticket, so every
// tile's predecessors are already resident), publishes its aggregate, then
// its first wave reads up to 64 predecessor descriptors at once until an
// inclusive prefix.  A descriptor packs (epoch:30 | flag:2 | value:32); the
// binning's epoch hides stale descriptors.  The wait is bounded: past
// kSpinLimit polls the binning is marked invalid instead.
constexpr long long kSpinLimit = 1ll << 24;

__device__ __forceinline__ unsigned long long scan_pack(unsigned epoch, unsigned flag, int v) {
  return ((unsigned long long)epoch << 34) | ((unsigned long long)flag << 32) | (unsigned)v;
}

__global__ void __launch_bounds__(kScanThreads) k_bl_scan(Work W, ScnState *ss) {
  __shared__ int s_tile, s_prefix;
  __shared__ int s_wave[kScanThreads / 64];
  DevState *st = W.st;
  if (ss->oob[0]) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      ss->valid = 0;
      st->status |= CD_STATUS_GRID_OVERFLOW;
    }
    return;
  }
  if (!ss->valid) return;
  const int C = ss->C;
  const unsigned epoch = st->scan_epoch;
  const int ntiles = (C + kScanTile - 1) / kScanTile;
  if (threadIdx.x == 0) s_tile = (int)atomicAdd(&st->scan_ticket, 1u);
  __syncthreads();
  const int tile = s_tile;
  if (tile >= ntiles) return;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int base = tile * kScanTile + t * kScanItems;
  int v[kScanItems];
  int sum = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int idx = base + k;
    v[k] = idx < C ? W.count[idx] : 0;
    sum += v[k];
  }
  int inc = sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(inc, o, 64);
    if (lane >= o) inc += u;
  }
  if (lane == 63) s_wave[wid] = inc;
  __syncthreads();
  int wave_off = 0, agg = 0;
#pragma unroll
  for (int w = 0; w < kScanThreads / 64; ++w) {
    if (w < wid) wave_off += s_wave[w];
    agg += s_wave[w];
  }
  const int excl = wave_off + inc - sum;
  unsigned long long *tiles = W.tiles;
  if (wid == 0) {
    if (lane == 0)
      __hip_atomic_store(&tiles[tile], scan_pack(epoch, tile == 0 ? 2u : 1u, agg),
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int prefix = 0;
    int p = tile - 1;
    long long spins = 0;
    bool failed = false;
    while (p >= 0) {
      const int idx = p - lane;
      unsigned flag = 2u;
      int val = 0;
      if (idx >= 0) {
        const unsigned long long d =
            __hip_atomic_load(&tiles[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        flag = ((unsigned)(d >> 34) == epoch) ? ((unsigned)(d >> 32) & 3u) : 0u;
        val = (int)(unsigned)d;
      }
      const unsigned long long incl = __ballot(flag == 2u);
      const unsigned long long bad = __ballot(flag == 0u);
      const int stop = incl ? (__ffsll((long long)incl) - 1) : 63;
      const unsigned long long need = (stop >= 63) ? ~0ull : ((2ull << stop) - 1ull);
      if (bad & need) {
        if (++spins > kSpinLimit) {
          failed = true;
          break;
        }
        continue;
      }
      int part = lane <= stop ? val : 0;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o, 64);
      prefix += part;
      if (incl) break;
      p -= 64;
    }
    if (lane == 0) {
      if (failed) {
        ss->valid = 0;
        st->status |= CD_STATUS_GRID_OVERFLOW;
      }
      if (tile != 0)
        __hip_atomic_store(&tiles[tile], scan_pack(epoch, 2u, prefix + agg), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      s_prefix = prefix;
    }
  }
  __syncthreads();
  int run = s_prefix + excl;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int idx = base + k;
    if (idx < C) W.offset[idx] = run;
    run += v[k];
  }
  if (tile == ntiles - 1 && t == kScanThreads - 1) W.offset[C] = s_prefix + agg;
}

__global__ void __launch_bounds__(256) k_bl_scatter(int n, Work W, const ScnState *ss) {
  if (!ss->valid) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  W.pidx[W.offset[W.cell_of[i]] + W.rank[i]] = i;
}

// Orders each cell's segment by agent index (stable counting sort).
__global__ void __launch_bounds__(256) k_bl_sort(Work W, const ScnState *ss) {
  if (!ss->valid) return;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ss->C) return;
  const int s0 = W.offset[c], s1 = W.offset[c + 1];
  for (int s = s0 + 1; s < s1; ++s) {
    const int key = W.pidx[s];
    int t = s - 1;
    while (t >= s0 && W.pidx[t] > key) {
      W.pidx[t + 1] = W.pidx[t];
      --t;
    }
    W.pidx[t + 1] = key;
  }
}

// ---------------------------------------------------------------------------
// Per-agent operators
// ---------------------------------------------------------------------------

// Fluctuation noise of agent i (core/motion/fluctuation.py:14-62): force
// (mass * sigma * |TN(0,3)|) * (cos phi, sin phi), phi ~ U(0, 2 pi); torque
// inertia * sigma_t * TN(-3, 3) when `torque`.  `i` is the agent's counter
// (its index within its scenario).
__device__ __forceinline__ void fluct_noise(uint64_t seed, uint64_t step, int i, double mass,
                                            double sig_f, bool torque, double inertia,
                                            double sig_t, double &fx, double &fy, double &m) {
  double ua, ub;
  draw2(seed, i, step, 0, 0, ua, ub);
  // (cos, sin)(2 pi ua) as sincospi(2 ua): exact argument, short reduction
  double sp, cp;
  sincospi(2.0 * ua, &sp, &cp);
  const double mm = mass * (sig_f * fabs(truncnorm3(seed, i, step, 1)));
  fx = mm * cp;
  fy = mm * sp;
  m = torque ? inertia * (sig_t * truncnorm3(seed, i, step, 2)) : 0.0;
}

template <typename T>
__global__ void __launch_bounds__(256) k_fluct(Agents<T> A, int model, uint64_t seed,
                                               uint64_t step) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  double fx, fy, m;
  fluct_noise(seed, step, i, A.mass[i], A.sig_f[i], model != 0, model ? A.inertia[i] : 0.0,
              model ? A.sig_t[i] : 0.0, fx, fy, m);
  V2<T> f = A.force[i];
  f.x += (T)fx;
  f.y += (T)fy;
  A.force[i] = f;
  if (model) A.torque[i] += m;
}

template <typename T>
__global__ void __launch_bounds__(256) k_nav(Agents<T> A, const double *maps, int64_t T_,
                                             int64_t rows, int64_t cols, double minx,
                                             double miny, double step, int mode) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  const double2 p = A.pos[i];
  double ex, ey;
  if (nav_lookup(p.x, p.y, A.target[i], maps, T_, rows, cols, minx, miny, step, mode, ex, ey)) {
    V2<T> e;
    e.x = (T)ex;
    e.y = (T)ey;
    A.e0[i] = e;
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_orient(Agents<T> A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  const V2<T> e = A.e0[i];
  A.phi0[i] = atan2((double)e.y, (double)e.x);  // orientation.py:21
}

// force_adjust (core/motion/adjusting.py:16-51)
template <typename T>
__device__ __forceinline__ void adjust_force(const Agents<T> &A, int i, V2<T> v, V2<T> e, T &fx,
                                             T &fy) {
  const T s = (T)(A.mass[i] / A.tau_adj[i]);
  const T v0 = (T)A.v0[i];
  fx = s * (v0 * e.x - v.x);
  fy = s * (v0 * e.y - v.y);
}

// torque_adjust (core/motion/adjusting.py:54-95)
template <typename T>
__device__ __forceinline__ double adjust_torque(const Agents<T> &A, int i) {
  const double s = A.inertia[i] / A.tau_rot[i];
  return s * ((wrap_to_pi(A.phi0[i] - A.phi[i]) / kPi) * A.w0[i] - A.omega[i]);
}

template <typename T>
__global__ void __launch_bounds__(256) k_adjust(Agents<T> A, int do_force, int do_torque) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  if (do_force) {
    T fx, fy;
    adjust_force(A, i, A.vel[i], A.e0[i], fx, fy);
    V2<T> f = A.force[i];
    f.x += fx;
    f.y += fy;
    A.force[i] = f;
  }
  if (do_torque) A.torque[i] += adjust_torque(A, i);
}

template <typename T>
__global__ void __launch_bounds__(256) k_reset(Agents<T> A, int model) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  V2<T> z;
  z.x = T(0);
  z.y = T(0);
  A.force[i] = z;
  if (model) A.torque[i] = 0.0;
}

// Agent-side inputs of the wall contacts (loaded once per agent).  reach2:
// the square of how far from the agent's centre its body extends (radius,
// or shoulder offset + the larger radius for three-circle agents), plus
// 1e-6 m; a segment farther than that from the centre cannot touch it.
template <typename T>
struct WallSelf {
  double r, rt, rsr, reach2;
  double2 ls, rs;
  T mu, kappa, damping;
};

template <typename T>
__device__ __forceinline__ WallSelf<T> wall_self(const Agents<T> &A, int model, int i, double2 p) {
  WallSelf<T> w;
  w.r = A.radius[i];
  w.ls = p;
  w.rs = p;
  w.rt = 0.0;
  w.rsr = 0.0;
  if (model) {
    w.ls = A.ls[i];
    w.rs = A.rs[i];
    w.rt = A.r_t[i];
    w.rsr = A.r_s[i];
  }
  w.mu = (T)A.mu[i];
  w.kappa = (T)A.kappa[i];
  w.damping = (T)A.damping[i];
  const double reach = (model ? A.r_ts[i] + fmax(w.rt, w.rsr) : w.r) + 1e-6;
  w.reach2 = reach * reach;
  return w;
}

// Wall contacts of an agent at p with velocity v over segments w_begin,
// w_begin + w_stride, ... (core/interactions.py:107-141, 169-186):
// add(force x, force y, torque) for each contact.
// `frm`: the segments' frames (seg_frame), or null to compute them here.
// `box`: the segments' bounding boxes {min x, min y, max x, max y}, or null;
// with boxes, a segment whose box lies farther than the agent's reach from
// its centre is skipped.  Every circle of the agent is then more than 1e-6 m
// from the segment, so the skipped contact term is exactly the zero that
// evaluating it would give.
template <typename T, class Add>
__device__ __forceinline__ void walls_each(const WallSelf<T> &me, int model, double2 p, V2<T> v,
                                           const double4 *seg, const double4 *frm, int W,
                                           Add &&add, int w_begin = 0, int w_stride = 1,
                                           const double4 *box = nullptr) {
  for (int w = w_begin; w < W; w += w_stride) {
    if (box) {
      const double4 b = box[w];
      const double dx = fmax(fmax(b.x - p.x, p.x - b.z), 0.0);
      const double dy = fmax(fmax(b.y - p.y, p.y - b.w), 0.0);
      if (dx * dx + dy * dy > me.reach2) continue;
    }
    const double4 s = seg[w];
    const double4 fr = frm ? frm[w] : seg_frame(s.x, s.y, s.z, s.w);
    double h, nx, ny, rmx = 0.0, rmy = 0.0;
    if (!model) {
      distance_circle_line(p.x, p.y, me.r, s.x, s.y, s.z, s.w, fr, h, nx, ny);
    } else {
      // distance_three_circle_line, core/distance.py:149-180
      const double px[3] = {p.x, me.ls.x, me.rs.x}, py[3] = {p.y, me.ls.y, me.rs.y};
      const double rr[3] = {me.rt, me.rsr, me.rsr};
      double h_min = __builtin_nan("");
      double bx = 0.0, by = 0.0, sx = 0.0, sy = 0.0, sr_ = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double hk, ex, ey;
        distance_circle_line(px[k], py[k], rr[k], s.x, s.y, s.z, s.w, fr, hk, ex, ey);
        if (hk < h_min || __builtin_isnan(h_min)) {
          h_min = hk;
          bx = ex;
          by = ey;
          sx = px[k];
          sy = py[k];
          sr_ = rr[k];
        }
      }
      h = h_min;
      nx = bx;
      ny = by;
      rmx = (sx - sr_ * nx) - px[0];
      rmy = (sy - sr_ * ny) - py[0];
    }
    if (h < 0.0) {
      T cx, cy;
      const T tnx = (T)nx, tny = (T)ny;
      force_contact((T)h, tnx, tny, v.x, v.y, tny, -tnx, me.mu, me.kappa, me.damping, cx, cy);
      add(cx, cy, model ? rmx * (double)cy - rmy * (double)cx : 0.0);
    }
  }
}

// The same for agent i, summed in segment order into fx, fy, tq.
template <typename T>
__device__ __forceinline__ void walls_agent(const Agents<T> &A, int model, int i, double2 p,
                                            V2<T> v, const double4 *seg, int W, T &fx, T &fy,
                                            double &tq) {
  walls_each(wall_self(A, model, i, p), model, p, v, seg, nullptr, W, [&](T cx, T cy, double m) {
    fx += cx;
    fy += cy;
    if (model) tq += m;
  });
}

constexpr int kSegTile = 512;

// AgentObstacleInteractions node: walls in the reference's order.
template <typename T>
__global__ void __launch_bounds__(256) k_walls(Agents<T> A, int model, const double4 *segs,
                                               int W) {
  __shared__ double4 sseg[kSegTile];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < A.n;
  double2 p;
  V2<T> v, f;
  double tq = 0.0;
  if (live) {
    p = A.pos[i];
    v = A.vel[i];
    f = A.force[i];
    if (model) tq = A.torque[i];
  }
  for (int w0 = 0; w0 < W; w0 += kSegTile) {
    const int cnt = min(kSegTile, W - w0);
    __syncthreads();
    for (int k = threadIdx.x; k < cnt; k += blockDim.x) sseg[k] = segs[w0 + k];
    __syncthreads();
    if (live) walls_agent(A, model, i, p, v, sseg, cnt, f.x, f.y, tq);
  }
  if (live) {
    A.force[i] = f;
    if (model) A.torque[i] = tq;
  }
}

// ---------------------------------------------------------------------------
// K_e: forces (owner-computes over the 3x3 cell stencil)
// ---------------------------------------------------------------------------
//
// L = 2, 4, 8, 16, 32 or 64 lanes serve one agent: the candidates of its 3 x 3 cell
// stencil, taken as one sequence, are split evenly over the lanes, and so
// are the wall segments.  (The agent's own terms were added by the step's
// first kernel.)  Pair and wall terms are
// summed exactly (Fix128), so the result does not depend on the order in
// which the binning inserted agents into cells or on how lanes and
// workgroups split the work; it is the correctly rounded sum of the terms up
// to the 2^-64 N grid.  Two kernels share this per-agent work:
//
//  * direct: workgroups take 16..64 agents in agent order; neighbour ids
//    come from the cells' slot lists and neighbour data is read by id.
//  * tiled: a workgroup takes a tile of tile_x * tile_y grid cells and
//    stages the tile plus a one-cell halo in LDS (cell counts -> offsets by
//    a block scan -> ids from the slot lists -> neighbour data by id), then
//    serves the agents of the tile's inner cells from LDS.  Each cell's data
//    is read from memory by at most four workgroups.  A region holding more
//    agents than the LDS capacity is processed cell by cell; a cell whose
//    3 x 3 region still does not fit is served from memory.
//
// Both give identical results; the tile shape is a tuning parameter.
// Members of a cell holding more agents than slots are found by scanning
// cell_of (exact, slow, flagged by SLOTS_FULL).


constexpr int kTileThreads = 256;
constexpr int kTileGroups = kTileThreads / 4;  // agents served at once
constexpr int kTileMaxDim = 16;
constexpr int kRegionMax = (kTileMaxDim + 2) * (kTileMaxDim + 2);
constexpr int kWallTile = 256;

// LDS capacity (agents of a region): nine cells of up to 128 (circular) or
// 56 (three-circle) agents fit, so a 3 x 3 region fits whenever no cell
// exceeds those counts.
template <int MODEL>
struct TileCap;
template <>
struct TileCap<0> {
  static constexpr int v = 1152;
};
template <>
struct TileCap<1> {
  static constexpr int v = 512;
};

template <typename T, int MODEL>
struct TileLds {
  static constexpr int CAP = TileCap<MODEL>::v;
  static constexpr int XC = MODEL ? CAP : 1;  // three-circle-only arrays
  int cnt[kRegionMax];
  int off[kRegionMax + 1];
  int own[kTileMaxDim + 1];
  int wsum[kTileThreads / 64];
  int id[CAP];
  int cell[CAP];  // region cell of each element
  double2 pos[CAP];
  V2<T> vel[CAP];
  double r[CAP];
  double2 ls[XC], rs[XC];
  double rt[XC], rsr[XC];
  double2 sho[XC];  // shoulder offsets
};

// Exclusive scan of one int per thread over the block; *total gets the sum.
__device__ __forceinline__ int block_scan_excl(int v, int *wsum, int &total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(inc, o, 64);
    if (lane >= o) inc += u;
  }
  if (lane == 63) wsum[wid] = inc;
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kTileThreads / 64; ++w) {
    const int s = wsum[w];
    if (w < wid) off += s;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return off + inc - v;
}

template <typename T>
__device__ __forceinline__ Owner<T> owner_of(const Agents<T> &A, int i) {
  Owner<T> o;
  o.neg_mk = (T)((-A.mass[i]) * A.k_soc[i]);
  o.tau_0 = (T)A.tau_0[i];
  o.mu = (T)A.mu[i];
  o.kappa = (T)A.kappa[i];
  o.damping = (T)A.damping[i];
  if constexpr (sizeof(T) == 4)
    o.rtau_0 = frcp_normal(o.tau_0);
  else
    o.rtau_0 = T(0);
  return o;
}

template <typename T>
__device__ __forceinline__ TCView tc_agent(const Agents<T> &A, int j) {
  TCView t;
  const double2 p = A.pos[j], l = A.ls[j], r = A.rs[j];
  t.px[0] = p.x;
  t.py[0] = p.y;
  t.px[1] = l.x;
  t.py[1] = l.y;
  t.px[2] = r.x;
  t.py[2] = r.y;
  const V2<T> v = A.vel[j];
  t.vx = (double)v.x;
  t.vy = (double)v.y;
  t.r_t = A.r_t[j];
  t.r_s = A.r_s[j];
  const double2 o = shoulder_offset(A.r_ts[j], A.phi[j]);
  t.ox = o.x;
  t.oy = o.y;
  return t;
}

// Where the force kernels read neighbour state: the agent arrays (operator
// path) or the snapshot the step's first kernel wrote (fused step, whose
// force kernel moves its agents while other agents still read them).
template <typename T>
struct Nbr {
  const double2 *pos;
  const V2<T> *vel;
  const double2 *ls, *rs;
  const double *phi;    // the agent arrays' orientation (operator path) ...
  const double2 *off;   // ... or the snapshot's shoulder offsets (fused step)
};

// A slot's neighbour record (circular agents, fused step): the data a
// candidate needs, written by the first kernel beside the slot's id, so the
// force kernel reads it at one address with no id hop.  The same values as
// the snapshot.  48 bytes in fp64, 32 in fp32 (the workspace holds 48 per
// slot).
template <typename T>
struct __attribute__((aligned(16))) NbrRec {
  double2 pos;
  V2<T> vel;
  double r;
};
static_assert(sizeof(NbrRec<double>) == 48 && sizeof(NbrRec<float>) == 32, "NbrRec layout");

template <typename T>
__host__ __device__ __forceinline__ Nbr<T> nbr_agents(const Agents<T> &A) {
  Nbr<T> N;
  N.pos = A.pos;
  N.vel = A.vel;
  N.ls = A.ls;
  N.rs = A.rs;
  N.phi = A.phi;
  N.off = nullptr;
  return N;
}

template <typename T>
__host__ __device__ __forceinline__ Nbr<T> nbr_snapshot(const Work &W) {
  Nbr<T> N;
  N.pos = W.snap_pos;
  N.vel = (const V2<T> *)W.snap_vel;
  N.ls = W.snap_ls;
  N.rs = W.snap_rs;
  N.phi = nullptr;
  N.off = W.snap_off;
  return N;
}

template <typename T>
__device__ __forceinline__ TCView tc_nbr(const Agents<T> &A, const Nbr<T> &N, int j) {
  TCView t;
  const double2 p = N.pos[j], l = N.ls[j], r = N.rs[j];
  t.px[0] = p.x;
  t.py[0] = p.y;
  t.px[1] = l.x;
  t.py[1] = l.y;
  t.px[2] = r.x;
  t.py[2] = r.y;
  const V2<T> v = N.vel[j];
  t.vx = (double)v.x;
  t.vy = (double)v.y;
  t.r_t = A.r_t[j];
  t.r_s = A.r_s[j];
  const double2 o = N.off ? N.off[j] : shoulder_offset(A.r_ts[j], N.phi[j]);
  t.ox = o.x;
  t.oy = o.y;
  return t;
}

template <typename T, int MODEL>
__device__ __forceinline__ TCView tc_lds(const TileLds<T, MODEL> &S, int e) {
  TCView t;
  const double2 p = S.pos[e], l = S.ls[e], r = S.rs[e];
  t.px[0] = p.x;
  t.py[0] = p.y;
  t.px[1] = l.x;
  t.py[1] = l.y;
  t.px[2] = r.x;
  t.py[2] = r.y;
  const V2<T> v = S.vel[e];
  t.vx = (double)v.x;
  t.vy = (double)v.y;
  t.r_t = S.rt[e];
  t.r_s = S.rsr[e];
  t.ox = S.sho[e].x;
  t.oy = S.sho[e].y;
  return t;
}

// Pair sums of one lane.  A float term pair (the fp32 mode) whose parts
// are both below 2^20 gathers its grid indices round(x 2^31) (Fix128::add
// of a float) in two 64-bit integers, folded into the exact sums after
// kFold such pairs and at the end (fold): an index is below 2^51 in
// magnitude, so kFold of them cannot overflow.  The sums equal those of
// adding every term to the Fix128 pair directly.
struct PairSum {
  static constexpr int kFold = 2048;
  Fix128 x, y, m;
  long long gx = 0, gy = 0;
  int n = 0;
  __device__ __forceinline__ void add(double px, double py) {
    x.add(px);
    y.add(py);
  }
  __device__ __forceinline__ void add(float px, float py) {
    if (__builtin_fabsf(px) < 0x1p20f && __builtin_fabsf(py) < 0x1p20f) {
      gx += Fix128::grid31(px);
      gy += Fix128::grid31(py);
      if (++n == kFold) fold();
    } else {
      x.add(px);
      y.add(py);
    }
  }
  __device__ __forceinline__ void fold() {
    x.add_grid31(gx);
    y.add_grid31(gy);
    gx = gy = 0;
    n = 0;
  }
};

// Circular: agent i's side of its pair with a neighbour at xj, vj, rj.
template <typename T>
__device__ __forceinline__ void pair_add_c(double2 xi, V2<T> vi, double ri, const Owner<T> &o,
                                           double2 xj, V2<T> vj, double rj, PairSum &ps) {
  T px, py;
  if (pair_circular<T>((T)(xi.x - xj.x), (T)(xi.y - xj.y), vi.x - vj.x, vi.y - vj.y,
                       (T)(ri + rj), o, px, py)) {
    ps.add(px, py);
  }
}

// Selects a or b field by field (one inlined pair function for both
// argument orders below).
__device__ __forceinline__ TCView tc_select(bool take_b, const TCView &a, const TCView &b) {
  TCView t;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    t.px[k] = take_b ? b.px[k] : a.px[k];
    t.py[k] = take_b ? b.py[k] : a.py[k];
  }
  t.vx = take_b ? b.vx : a.vx;
  t.vy = take_b ? b.vy : a.vy;
  t.r_t = take_b ? b.r_t : a.r_t;
  t.r_s = take_b ? b.r_s : a.r_s;
  t.ox = take_b ? b.ox : a.ox;
  t.oy = take_b ? b.oy : a.oy;
  return t;
}

// Three-circle: the reference evaluates the pair as interaction(min, max).
__device__ __forceinline__ void pair_add_tc(int i, int j, const TCView &me, const TCView &other,
                                            const Owner<double> &o, PairSum &ps) {
  double px, py, m;
  const bool sb = !(i < j);
  const bool hit = pair_three_circle(tc_select(sb, me, other), tc_select(sb, other, me), sb, o,
                                     px, py, m);
  if (hit) {
    ps.x.add(px);
    ps.y.add(py);
    ps.m.add(m);
  }
}

// The same in the fp32 mode (float force terms, double torque).
__device__ __forceinline__ void pair_add_tc(int i, int j, const TCView &me, const TCView &other,
                                            const Owner<float> &o, PairSum &ps) {
  float px, py;
  double m;
  const bool sb = !(i < j);
  const bool hit = pair_three_circle_f32(tc_select(sb, me, other), tc_select(sb, other, me), sb,
                                         o, px, py, m);
  if (hit) {
    ps.add(px, py);
    ps.m.add(m);
  }
}

// Own (per-agent) terms of the fused step in the reference's post-order:
// fluctuation, navigation (e0), orientation (phi0), adjusting; force
// (torque) += fluctuation + adjusting.  own_load issues the agent's loads,
// own_nav the navigation map's (it needs the position and target), and
// own_apply computes and stores, so a kernel can put independent work
// between them; own_apply computes the noise before it first uses the map.
template <typename T>
struct OwnIn {
  V2<T> f, vel, e;
  double tq, mass, sig_f, inertia, sig_t, tau_adj, v0;
  int64_t target;
};

template <typename T, int MODEL>
__device__ __forceinline__ OwnIn<T> own_load(const Agents<T> &A, const FusedArgs &fa, int i,
                                             double2 pos) {
  OwnIn<T> o;
  o.f = A.force[i];
  o.tq = MODEL ? A.torque[i] : 0.0;
  o.vel = A.vel[i];
  o.mass = A.mass[i];
  o.sig_f = o.inertia = o.sig_t = o.tau_adj = o.v0 = 0.0;
  if (fa.flags & CD_USE_FLUCT) {
    o.sig_f = A.sig_f[i];
    if (MODEL) {
      o.inertia = A.inertia[i];
      o.sig_t = A.sig_t[i];
    }
  }
  o.e = A.e0[i];
  o.target = (fa.flags & CD_USE_NAV) ? A.target[i] : -1;
  if (fa.flags & CD_USE_ADJUST) {
    o.tau_adj = A.tau_adj[i];
    o.v0 = A.v0[i];
  }
  return o;
}

template <typename T>
__device__ __forceinline__ NavFetch own_nav(const FusedArgs &fa, double2 pos, const OwnIn<T> &o) {
  return nav_fetch(pos.x, pos.y, o.target, fa.maps, fa.nav_T, fa.nav_rows, fa.nav_cols,
                   fa.nav_minx, fa.nav_miny, fa.nav_step, fa.nav_mode);
}

// `local` is the agent's index within its scenario (the noise counter).
template <typename T, int MODEL>
__device__ __forceinline__ void own_apply(const Agents<T> &A, const FusedArgs &fa, uint64_t seed,
                                          uint64_t step, int i, int local, const OwnIn<T> &o,
                                          const NavFetch &nf) {
  V2<T> f = o.f;
  double tq = o.tq;
  if (fa.flags & CD_USE_FLUCT) {
    double nfx, nfy, nm;
    fluct_noise(seed, step, local, o.mass, o.sig_f, MODEL != 0, o.inertia, o.sig_t, nfx, nfy, nm);
    f.x += (T)nfx;
    f.y += (T)nfy;
    if (MODEL) tq += nm;
  }
  V2<T> e = o.e;
  double ex, ey;
  if ((fa.flags & CD_USE_NAV) && nav_finish(nf, ex, ey)) {
    e.x = (T)ex;
    e.y = (T)ey;
    A.e0[i] = e;
  }
  if (fa.flags & CD_USE_ADJUST) {
    if (MODEL) {
      // orient_towards_target_direction (core/steering/orientation.py:17-21),
      // or, with CD_USE_ORIENT_VELOCITY, toward the velocity the step starts from
      A.phi0[i] = (fa.flags & CD_USE_ORIENT_VELOCITY) ? atan2((double)o.vel.y, (double)o.vel.x)
                                                       : atan2((double)e.y, (double)e.x);
    }
    // force_adjust (core/motion/adjusting.py:16-51), as adjust_force
    const T s = (T)(o.mass / o.tau_adj);
    const T v0 = (T)o.v0;
    f.x += s * (v0 * e.x - o.vel.x);
    f.y += s * (v0 * e.y - o.vel.y);
    if (MODEL) tq += adjust_torque(A, i);
  }
  A.force[i] = f;
  if (MODEL) A.torque[i] = tq;
}

// K_a of the fused step (step index s = the counter, parity p = s & 1).
// Per scenario: read the reduction of the state the step starts from
// (acc[p], merged by the previous step's force kernel or by the
// preparation), publish the grid, dt and time_tot, reset acc[1 - p] and
// the other buffer's out-of-grid flag for this step's force kernel.  Per
// agent: snapshot the state (neighbour data of the force kernel), bin it
// into count buffer p (bin_scenario's contract), add the step's own terms.
// Count buffer 1 - p, filled by the previous step, is cleared here.  The
// own terms need neither grid nor neighbours: the first agent's inputs are
// loaded before the reduction is read, and its noise is computed while its
// slot atomic and navigation map loads are in flight.  The counter is read
// here only; the force kernel advances it.
// SPLIT (fluctuation or navigation on): the first half of each block bins
// its agents (snapshot, slot atomic, slot list and record) and the second
// half adds the same agents' own terms (noise, navigation, adjusting).  The
// two parts share no data, so the halves run side by side, and a lone wave
// per SIMD (one scenario: 128-thread blocks) no longer runs both dependent
// chains one after the other.
template <typename T, int MODEL, bool SPLIT, bool ONE>
__global__ void __launch_bounds__(256) k_step_begin(Agents<T> A, Work W, Batch B, FusedArgs fa,
                                                    double cell, int binning) {
  const int2 sk = block_f_t<ONE>(B);
  const ScnDesc d = desc_t<ONE>(B, sk.x);
  ScnState *ss = &B.sst[sk.x];
  const uint64_t step = W.st->step;
  const int p = (int)(step & 1ull);
  int *count = W.count + (size_t)p * W.cap;
  // blocks of the scenario: its block map's, or the whole grid for a single
  // scenario (launched as 64-thread blocks, one wave each, spread over all
  // CUs; 128-thread blocks with SPLIT)
  const int nblk = ONE ? (int)gridDim.x : d.nf;
  const int per = SPLIT ? (int)blockDim.x / 2 : (int)blockDim.x;  // agents of a block at once
  const bool own_wave = SPLIT && (int)threadIdx.x >= per;
  const bool bin_wave = !own_wave;
  const int lane = own_wave ? (int)threadIdx.x - per : (int)threadIdx.x;
  const int stride = nblk * per;
  const int i0 = d.a0 + sk.y * per + lane;
  double2 p0 = make_double2(0.0, 0.0);
  OwnIn<T> o0;
  V2<T> v0;
  if (i0 < d.a1) {
    p0 = A.pos[i0];
    if (!SPLIT || own_wave) o0 = own_load<T, MODEL>(A, fa, i0, p0);
    if (SPLIT && bin_wave) v0 = A.vel[i0];
  }
  {  // counts of the previous step's binning (every thread of the block)
    int *other = W.count + (size_t)(1 - p) * W.cap + d.c0;
    const int clear = ss->C_clear[1 - p];
    for (int c = sk.y * (int)blockDim.x + (int)threadIdx.x; c < clear; c += nblk * (int)blockDim.x)
      other[c] = 0;
  }
  if (sk.y == 0)
    for (int sub = threadIdx.x; sub < kAccSub; sub += blockDim.x)
      acc_reset(acc_at(W, sk.x, 1 - p, sub));
  // an agent's own terms (map loads issued before the noise is computed)
  auto own = [&](int i, double2 pos, const OwnIn<T> &o) {
    NavFetch nf;
    nf.ok = false;
    if (fa.flags & CD_USE_NAV) nf = own_nav(fa, pos, o);
    own_apply<T, MODEL>(A, fa, d.seed, step, i, i - d.a0, o, nf);
  };
  if (own_wave) {  // needs no grid
    if (i0 < d.a1) own(i0, p0, o0);
    for (int i = i0 + stride; i < d.a1; i += stride) {
      const double2 pos = A.pos[i];
      own(i, pos, own_load<T, MODEL>(A, fa, i, pos));
    }
    return;
  }
  const Red r = acc_read(W, sk.x, p);
  const Grid g = grid_of(r, cell, d.ccap);
  if (sk.y == 0 && threadIdx.x == 0) {
    publish_grid(W, ss, r, g, binning != 0, d.a1 > d.a0, p);
    if (!binning) ss->C_clear[p] = 0;
    ss->oob[1 - p] = 0;
    ss->par = p;
    const double dt = adaptive_dt(r.v, d.dt_min, d.dt_max);
    ss->dt = dt;
    ss->time_tot += dt;
  }
  const bool bin = binning && g.ok;
  // steps that rebuild the force kernel's processing order keep the slot ranks
  const bool ranks = W.order_every > 0 && step % (uint64_t)W.order_every == 0ull;
  V2<T> *snap_vel = (V2<T> *)W.snap_vel;
  // an agent's snapshot and slot atomic (bin_begin), then its slot list
  // entry and neighbour record (bin_end)
  auto bin_begin = [&](int i, double2 pos, V2<T> vel, int &c, int &k, double &ri) {
    ri = (MODEL == 0 && bin) ? A.radius[i] : 0.0;  // for the neighbour record
    W.snap_pos[i] = pos;
    snap_vel[i] = vel;
    if (MODEL) {
      W.snap_ls[i] = A.ls[i];
      W.snap_rs[i] = A.rs[i];
      W.snap_off[i] = shoulder_offset(A.r_ts[i], A.phi[i]);
    }
    c = -1;
    k = 0;
    if (bin) {
      const int lc = cell_of_pos(pos, g, cell);
      if (lc < 0) {  // positions changed after the bounding box was reduced
        atomicOr(&ss->oob[p], 1);
        // the force kernel skips this scenario's pairs for the step: report it
        atomicOr((unsigned long long *)&W.st->status,
                 (unsigned long long)CD_STATUS_GRID_OVERFLOW);
        W.cell_of[i] = -1;
      } else {
        c = d.c0 + lc;
        W.cell_of[i] = c;
        k = atomicAdd(&count[c], 1);
        // the slot: for the processing order, and (circular agents) for the
        // force kernel, which skips the agent's own record by it
        if (ranks || MODEL == 0) W.rank[i] = k;
      }
    }
  };
  auto bin_end = [&](int i, double2 pos, V2<T> vel, int c, int k, double ri) {
    if (c < 0) return;
    if (k < W.K) {
      W.ell_ids[c * W.K + k] = i;
      if constexpr (MODEL == 0) {
        NbrRec<T> rr;
        rr.pos = pos;
        rr.vel = vel;
        rr.r = ri;
        ((NbrRec<T> *)W.rec)[(size_t)c * W.K + k] = rr;
      }
    } else if (k == W.K) {  // the cell's list is full: its members are found by scanning cell_of
      atomicOr((unsigned long long *)&W.st->status, (unsigned long long)CD_STATUS_SLOTS_FULL);
    }
  };
  if constexpr (SPLIT) {
    auto serve = [&](int i, double2 pos, V2<T> vel) {
      int c, k;
      double ri;
      bin_begin(i, pos, vel, c, k, ri);
      bin_end(i, pos, vel, c, k, ri);
    };
    if (i0 < d.a1) serve(i0, p0, v0);
    for (int i = i0 + stride; i < d.a1; i += stride) serve(i, A.pos[i], A.vel[i]);
  } else {
    // the own terms between the slot atomic and its use
    auto serve = [&](int i, double2 pos, const OwnIn<T> &o) {
      int c, k;
      double ri;
      bin_begin(i, pos, o.vel, c, k, ri);
      own(i, pos, o);
      bin_end(i, pos, o.vel, c, k, ri);
    };
    if (i0 < d.a1) serve(i0, p0, o0);
    for (int i = i0 + stride; i < d.a1; i += stride) {
      const double2 pos = A.pos[i];
      serve(i, pos, own_load<T, MODEL>(A, fa, i, pos));
    }
  }
}

// Processing order of the agent-parallel force kernel (and of the adjoint
// kernel, which follows the forward steps' order): each scenario's agents
// by cell, a counting sort of the step's binning (exclusive scan of
// the scenario's cell counts, then each agent at its cell's offset plus its
// slot rank).  A wave's agents then share their neighbour cells, so the
// lanes that walk the same part of the stencil load the same candidates.
// One block per scenario, after K_a, on steps whose counter is a multiple
// of order_every.  A scenario whose binning is not valid, or whose counts do
// not add up to its agents, keeps its previous order: the order always
// holds each scenario's agents once.
constexpr int kOrderThreads = 1024;
__global__ void __launch_bounds__(kOrderThreads) k_order_build(Work W, Batch B) {
  __shared__ int s_wave[kOrderThreads / 64];
  if (W.st->step % (uint64_t)W.order_every != 0ull) return;
  const ScnDesc d = desc_of(B, (int)blockIdx.x);
  const ScnState *ss = &B.sst[blockIdx.x];
  const int p = ss->par;
  const int n = d.a1 - d.a0;
  if (!ss->valid || ss->oob[p] || n <= 0) return;
  const int C = ss->C;
  const int *count = W.count + (size_t)p * W.cap + d.c0;
  int *off = W.order_off + d.c0;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int per = (C + kOrderThreads - 1) / kOrderThreads;
  const int k0 = min(C, t * per), k1 = min(C, k0 + per);
  int sum = 0;
  for (int k = k0; k < k1; ++k) sum += count[k];
  int inc = sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(inc, o, 64);
    if (lane >= o) inc += u;
  }
  if (lane == 63) s_wave[wid] = inc;
  __syncthreads();
  int before = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kOrderThreads / 64; ++w) {
    if (w < wid) before += s_wave[w];
    total += s_wave[w];
  }
  if (total != n) return;  // the same for every thread of the block
  int run = before + inc - sum;
  for (int k = k0; k < k1; ++k) {
    off[k] = run;
    run += count[k];
  }
  __syncthreads();  // offsets visible to the block
  for (int i = d.a0 + t; i < d.a1; i += kOrderThreads) {
    const int k = W.cell_of[i] - d.c0;
    if (k < 0 || k >= C) continue;
    const int e = off[k] + W.rank[i];
    if (e >= 0 && e < n) W.order[d.a0 + e] = i;
  }
}

// Identity processing order (preparation).
__global__ void __launch_bounds__(256) k_order_identity(int *order, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    order[i] = i;
}

// Neighbour sources.  each(q, L, self, fn) visits part q (of L equal
// parts) of the candidates of the agent's 3 x 3 stencil, taken as one
// sequence: the cells of rows dx = -1, 0, +1, each row's cells dy = -1, 0,
// +1.  Parts may visit candidates in any order (the pair sums are exact).
// From the LDS elements of a staged region (each row's three cells are
// consecutive region cells, so their elements are one contiguous range) ...
template <typename T, int MODEL>
struct LdsRow {
  const TileLds<T, MODEL> &S;
  int RH, kx, ky;  // region column length and the agent's region cell
  template <class F>
  __device__ __forceinline__ void each(int q, int L, int self, F &&fn) const {
    int st[3], pre[4];
    pre[0] = 0;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const int kk = (kx + r - 1) * RH + (ky - 1);
      st[r] = S.off[kk];
      pre[r + 1] = pre[r] + (S.off[kk + 3] - st[r]);
    }
    const int tot = pre[3];
    const int lg = __builtin_ctz(L);  // L is a power of two
    const int b0 = (tot * q) >> lg, b1 = (tot * (q + 1)) >> lg;
    for (int idx = b0; idx < b1; ++idx) {
      const int r = (idx >= pre[1] ? 1 : 0) + (idx >= pre[2] ? 1 : 0);
      const int e = (r == 0 ? st[0] : (r == 1 ? st[1] - pre[1] : st[2] - pre[2])) + idx;
      const int j = S.id[e];
      if (j == self) continue;
      if constexpr (MODEL == 0)
        fn(j, S.pos[e], S.vel[e], S.r[e]);
      else
        fn(j, tc_lds<T, MODEL>(S, e));
    }
  }
};

// ... or from the grid cells in memory: slot lists, or a scan of cell_of
// for a cell holding more agents than slots.
template <typename T, int MODEL>
struct MemRow {
  const Agents<T> &A;
  const Work &W;
  const int *count;  // the count buffer of this binning
  Nbr<T> N;
  int c0, nx, ny, lx, ly, a0, a1;
  // circular agents of the fused step: the slots' neighbour records, and the
  // agent's own slot in its cell (the stencil's centre cell); else null
  const NbrRec<T> *rec;
  int self_slot;
  template <class F>
  __device__ __forceinline__ void visit(int self, int j, F &fn) const {
    if (j == self) return;
    if constexpr (MODEL == 0)
      fn(j, N.pos[j], N.vel[j], A.radius[j]);
    else
      fn(j, tc_nbr(A, N, j));
  }
  // The nine counts are loaded together.  A part's candidates are walked
  // in batches of four whose id and data loads are issued together, so a
  // lane waits for memory once per batch rather than once per candidate.
  template <class F>
  __device__ __forceinline__ void each(int q, int L, int self, F &&fn) const {
    int gr[3], n[9];  // flat index of each row's cell dy = -1; counts
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const int ox = lx + r - 1;
      const bool okx = ox >= 0 && ox < nx;
      gr[r] = c0 + ox * ny + ly - 1;
      n[3 * r] = (okx && ly - 1 >= 0) ? count[gr[r]] : 0;
      n[3 * r + 1] = okx ? count[gr[r] + 1] : 0;
      n[3 * r + 2] = (okx && ly + 1 < ny) ? count[gr[r] + 2] : 0;
    }
    int pre[10];
    pre[0] = 0;
    bool over = false;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      pre[k + 1] = pre[k] + n[k];
      over = over || n[k] > W.K;
    }
    const int tot = pre[9];
    // floor(tot q / L) by a shift (L is a power of two), in 64 bits: the
    // counts of overfull cells are not bounded by the slots
    const int lg = __builtin_ctz(L);
    const int b0 = (int)(((long long)tot * q) >> lg);
    const int b1 = (int)(((long long)tot * (q + 1)) >> lg);
    if (!over) {
      // slot-list position of candidate idx in cell k: so[k] + idx (empty
      // cells, which hold no candidate, keep 0; every real cell * K is below
      // 2^31 by the workspace's size check)
      int so[9];
#pragma unroll
      for (int k = 0; k < 9; ++k)
        so[k] = n[k] ? (gr[k / 3] + k % 3) * W.K - pre[k] : 0;
      if constexpr (MODEL == 0) {
        if (rec) {
          // neighbour records: one load per candidate and no id hop; the
          // agent's own record (centre cell k = 4, its slot) is skipped by
          // its candidate index
          const int sidx = self_slot >= 0 ? pre[4] + self_slot : -1;
          // batches of kRecBatch candidates whose loads are issued together:
          // four fp32 records, two fp64 ones (a batch of four fp64 records
          // holds 40 registers, which made the fp64 kernel spill)
          constexpr int kRecBatch = sizeof(T) == 8 ? 2 : 4;
          for (int b = b0; b < b1; b += kRecBatch) {
            int s[kRecBatch];
#pragma unroll
            for (int u = 0; u < kRecBatch; ++u) {
              const int idx = b + u;
              s[u] = -1;
              if (idx < b1 && idx != sidx) {
                int o = so[0];
#pragma unroll
                for (int k = 1; k < 9; ++k)
                  if (idx >= pre[k]) o = so[k];
                s[u] = o + idx;
              }
            }
            NbrRec<T> rr[kRecBatch];
#pragma unroll
            for (int u = 0; u < kRecBatch; ++u)
              if (s[u] >= 0) rr[u] = rec[s[u]];
#pragma unroll
            for (int u = 0; u < kRecBatch; ++u)
              if (s[u] >= 0) fn(-1, rr[u].pos, rr[u].vel, rr[u].r);
            // The batch's loads have all been consumed; waiting for them
            // here says so to the compiler, which otherwise holds each
            // record load of the next batch until the previous one arrives
            // (their registers might, for all it knows, still be filling).
            // s_waitcnt vmcnt(0), expcnt and lgkmcnt left at their maxima.
            __builtin_amdgcn_s_waitcnt(0x0F70);
          }
          return;
        }
      }
      // batches as for the records: four candidates, two fp64 circular ones
      constexpr int kIdBatch = (MODEL == 0 && sizeof(T) == 8) ? 2 : 4;
      for (int b = b0; b < b1; b += kIdBatch) {
        int j[kIdBatch];
#pragma unroll
        for (int u = 0; u < kIdBatch; ++u) {
          const int idx = b + u;
          j[u] = -1;
          if (idx < b1) {
            // the cell holding candidate idx: the last k with pre[k] <= idx
            int o = so[0];
#pragma unroll
            for (int k = 1; k < 9; ++k)
              if (idx >= pre[k]) o = so[k];
            j[u] = W.ell_ids[o + idx];
          }
        }
        if constexpr (MODEL == 0) {
          double2 xj[kIdBatch];
          V2<T> vj[kIdBatch];
          double rj[kIdBatch];
#pragma unroll
          for (int u = 0; u < kIdBatch; ++u) {
            if (j[u] >= 0) {
              xj[u] = N.pos[j[u]];
              vj[u] = N.vel[j[u]];
              rj[u] = A.radius[j[u]];
            }
          }
#pragma unroll
          for (int u = 0; u < kIdBatch; ++u)
            if (j[u] >= 0 && j[u] != self) fn(j[u], xj[u], vj[u], rj[u]);
          __builtin_amdgcn_s_waitcnt(0x0F70);  // as for the records
        } else {
          // one copy of the three-circle pair code for the batch (the id
          // picked by selects, not by indexing): an unrolled batch held
          // eight inlined copies, more code than the instruction cache
#pragma unroll 1
          for (int u = 0; u < kIdBatch; ++u) {
            int ju = j[0];
#pragma unroll
            for (int w = 1; w < kIdBatch; ++w) ju = u == w ? j[w] : ju;
            if (ju >= 0) visit(self, ju, fn);
          }
          __builtin_amdgcn_s_waitcnt(0x0F70);  // as for the records
        }
      }
      return;
    }
    // a cell holding more agents than slots: members found by scanning
    // cell_of, the candidates numbered cell by cell (not unrolled: one copy
    // of the pair code)
#pragma unroll 1
    for (int k = 0; k < 9; ++k) {
      const int cnt = n[k];
      const int s0 = max(0, b0 - pre[k]), s1 = min(cnt, b1 - pre[k]);
      if (s0 >= s1) continue;
      const int gc = gr[k / 3] + k % 3;
      if (cnt <= W.K) {
        const int *ids = W.ell_ids + (size_t)gc * W.K;
        for (int s2 = s0; s2 < s1; ++s2) visit(self, ids[s2], fn);
      } else {
        int r = 0;
        for (int jx = a0; jx < a1 && r < s1; ++jx) {
          if (W.cell_of[jx] != gc) continue;
          if (r >= s0) visit(self, jx, fn);
          ++r;
        }
      }
    }
  }
};

// Stages cnt wall segments, their frames (seg_frame) and bounding boxes in
// LDS: sseg holds 3 kWallTile entries, the segments, then the frames, then
// the boxes.
__device__ __forceinline__ void stage_segs(const double4 *src, int cnt, double4 *sseg) {
  for (int k = threadIdx.x; k < cnt; k += blockDim.x) {
    const double4 s = src[k];
    sseg[k] = s;
    sseg[kWallTile + k] = seg_frame(s.x, s.y, s.z, s.w);
    sseg[2 * kWallTile + k] =
        make_double4(fmin(s.x, s.z), fmin(s.y, s.w), fmax(s.x, s.z), fmax(s.y, s.w));
  }
}

// Stages the scenario's first kWallTile wall segments (before the per-agent
// work, so the loads overlap it).
__device__ __forceinline__ void stage_walls(const FusedArgs &fa, const ScnDesc &d, double4 *sseg,
                                            bool fused) {
  if (!fused || !(fa.flags & CD_USE_OBSTACLE) || fa.wall_cells) return;
  stage_segs(fa.segs + d.w0, min(kWallTile, d.w1 - d.w0), sseg);
}

__device__ __forceinline__ Fix128 shfl_xor(const Fix128 &v, int mask) {
  Fix128 t;
  t.hi = __shfl_xor(v.hi, mask, 64);
  t.lo = __shfl_xor(v.lo, mask, 64);
  t.bad = __shfl_xor((int)v.bad, mask, 64) != 0;
  return t;
}

// translational_verlet (core/integrator.py:164-176), rotational_verlet
// (:179-193), shoulders (simulation/agents.py:473-486) and Reset
// (simulation/logic.py:59-64) of agent i, given its force f (torque tq),
// position x, velocity vel, previous force fp and mass m.  Writes position,
// velocity, previous force and the force (zero after Reset, else f).
template <typename T>
__device__ __forceinline__ void verlet_with(const Agents<T> &A, int model, int i, double dt,
                                            bool do_reset, V2<T> f, double tq, double2 x,
                                            V2<T> vel, V2<T> fp, double m, double2 &x_out,
                                            V2<T> &v_out) {
  const T tm = (T)m;
  const T aox = fp.x / tm, aoy = fp.y / tm;
  const T anx = f.x / tm, any = f.y / tm;
  vel.x += (aox + anx) / T(2) * (T)dt;
  vel.y += (aoy + any) / T(2) * (T)dt;
  if (sizeof(T) == 8) {
    x.x += (double)vel.x * dt + (double)anx / 2.0 * (dt * dt);
    x.y += (double)vel.y * dt + (double)any / 2.0 * (dt * dt);
  } else {
    // fp32 mode: fp64 position accumulation of the fp32 increments
    x.x += (double)vel.x * dt + (double)(anx / T(2)) * (dt * dt);
    x.y += (double)vel.y * dt + (double)(any / T(2)) * (dt * dt);
  }
  A.force_prev[i] = f;
  A.vel[i] = vel;
  A.pos[i] = x;
  V2<T> fz = f;
  if (do_reset) {
    fz.x = T(0);
    fz.y = T(0);
  }
  A.force[i] = fz;
  if (model) {
    const double I = A.inertia[i];
    const double ao = A.torque_prev[i] / I;
    const double an = tq / I;
    A.torque_prev[i] = tq;
    const double w = A.omega[i] + (ao + an) / 2.0 * dt;
    double ph = A.phi[i] + (w * dt + an / 2.0 * (dt * dt));
    ph = wrap_to_pi(ph);
    A.omega[i] = w;
    A.phi[i] = ph;
    const double rts = A.r_ts[i];
    const double ox = sin(ph) * rts, oy = (-cos(ph)) * rts;
    A.ls[i] = make_double2(x.x - ox, x.y - oy);
    A.rs[i] = make_double2(x.x + ox, x.y + oy);
    A.torque[i] = do_reset ? 0.0 : tq;
  }
  x_out = x;
  v_out = vel;
}

// Agent-side inputs of the force pass, loaded together before any
// neighbour is visited.
template <typename T>
struct Self {
  Owner<T> o;
  WallSelf<T> w;
  V2<T> f;    // force after the step's own terms
  double tq;  // torque after them
};

template <typename T, int MODEL>
__device__ __forceinline__ Self<T> self_of(const Agents<T> &A, int i, double2 pos) {
  Self<T> s;
  s.o = owner_of(A, i);
  s.w = wall_self(A, MODEL, i, pos);
  s.f = A.force[i];
  s.tq = MODEL ? A.torque[i] : 0.0;
  return s;
}

// Pair and wall forces of agent i, served by L lanes (L = 2 to 64; q =
// lane in the group): lane q takes part q of the stencil's candidates
// (when `pairs`) and every L-th wall from q (fused step with walls).  Pair
// and wall terms are summed exactly, so results do not depend on L.  Every
// lane of the block calls this (have = false for idle groups), so the wall
// staging barriers and the shuffles are uniform; the first kWallTile
// segments are already staged.  Every lane of the group returns force
// (torque) = self.f (self.tq) + (pairs + walls): the pair and wall terms
// are one exact sum, rounded once.
template <typename T, int MODEL, bool FUSED, class Rows>
__device__ __forceinline__ void agent_forces(const FusedArgs &fa, const ScnDesc &d,
                                             double4 *sseg, int L, bool have, bool pairs, int i,
                                             double2 pos, V2<T> vel, const TCView &me,
                                             const Self<T> &self, const Rows &rows, V2<T> &f,
                                             double &tq) {
  const int q = threadIdx.x & (L - 1);
  PairSum ps;  // pairs
  if (have && pairs) {
    if constexpr (MODEL == 0) {
      rows.each(q, L, i, [&](int, double2 xj, V2<T> vj, double rj) {
        pair_add_c<T>(pos, vel, self.w.r, self.o, xj, vj, rj, ps);
      });
    } else {
      rows.each(q, L, i, [&](int j, const TCView &other) {
        pair_add_tc(i, j, me, other, self.o, ps);
      });
    }
  }
  const bool walls = FUSED && (fa.flags & CD_USE_OBSTACLE);
  if (walls) {  // wall contacts join the same exact sums
    auto add = [&](T cx, T cy, double mm) {
      ps.add(cx, cy);
      if (MODEL) ps.m.add(mm);
    };
    if (fa.wall_cells) {
      // binned walls: only the segments listed for the agent's wall cell
      // (a superset of those within its reach; outside the grid, none)
      if (have) {
        const double qx = (pos.x - fa.wall_minx) / fa.wall_cell;
        const double qy = (pos.y - fa.wall_miny) / fa.wall_cell;
        if (qx >= 0.0 && qy >= 0.0 && qx < (double)fa.wall_nx && qy < (double)fa.wall_ny) {
          const int c = (int)qx * fa.wall_ny + (int)qy;
          const int k0 = fa.wall_cells[c], k1 = fa.wall_cells[c + 1];
          const int *ids = fa.wall_cells + (fa.wall_nx * fa.wall_ny + 1);
          for (int k = k0 + q; k < k1; k += L) {
            const double4 sg = fa.segs[d.w0 + ids[k]];
            const double4 bx = make_double4(fmin(sg.x, sg.z), fmin(sg.y, sg.w),
                                            fmax(sg.x, sg.z), fmax(sg.y, sg.w));
            walls_each(self.w, MODEL, pos, vel, &sg, nullptr, 1, add, 0, 1, &bx);
          }
        }
      }
    } else {
      const int nw = d.w1 - d.w0;
      const double4 *sfrm = sseg + kWallTile, *sbox = sseg + 2 * kWallTile;
      if (have)
        walls_each(self.w, MODEL, pos, vel, sseg, sfrm, min(kWallTile, nw), add, q, L, sbox);
      for (int w0 = kWallTile; w0 < nw; w0 += kWallTile) {  // more segments than staged
        const int cnt = min(kWallTile, nw - w0);
        __syncthreads();
        stage_segs(fa.segs + d.w0 + w0, cnt, sseg);
        __syncthreads();
        if (have) walls_each(self.w, MODEL, pos, vel, sseg, sfrm, cnt, add, q, L, sbox);
      }
      if (nw > kWallTile) {  // restore the first tile for the next agents
        __syncthreads();
        stage_walls(fa, d, sseg, true);
        __syncthreads();
      }
    }
  }
  if constexpr (sizeof(T) == 4) ps.fold();
  // exact sums over the L lanes (butterfly: every lane ends with the totals)
  for (int mask = 1; mask < L; mask <<= 1) {
    ps.x.add(shfl_xor(ps.x, mask));
    ps.y.add(shfl_xor(ps.y, mask));
    if (MODEL) ps.m.add(shfl_xor(ps.m, mask));
  }
  f = self.f;
  tq = self.tq;
  f.x += (T)ps.x.value();
  f.y += (T)ps.y.value();
  if (MODEL) tq += ps.m.value();
}

// Inputs of the integrator the force kernels of the fused step load early.
template <typename T>
struct Tail {
  V2<T> fp;
  double m, v0;
};

template <typename T>
__device__ __forceinline__ Tail<T> tail_of(const Agents<T> &A, int i) {
  Tail<T> t;
  t.fp = A.force_prev[i];
  t.m = A.mass[i];
  t.v0 = A.v0[i];
  return t;
}

// After an agent's force pass, on its lead lane: the fused step integrates
// it (Verlet and Reset) and adds its new state to the block's reduction r;
// the AgentAgentInteractions node stores the force.
template <typename T, int MODEL, bool FUSED>
__device__ __forceinline__ void agent_finish(const Agents<T> &A, const FusedArgs &fa, int i,
                                             bool lead, V2<T> f, double tq, double2 pos,
                                             V2<T> vel, const Tail<T> &t, double dt, int reset,
                                             Red &r) {
  if (!lead) return;
  if constexpr (FUSED) {
    double2 x;
    V2<T> v;
    verlet_with(A, MODEL, i, dt, reset != 0, f, tq, pos, vel, t.fp, t.m, x, v);
    r.add_agent(x, dhypot((double)v.x, (double)v.y), t.v0);
    if (fa.tape_x) fa.tape_x[i] = x;
    if (fa.tape_v) ((V2<T> *)fa.tape_v)[i] = v;
  } else {
    A.force[i] = f;
    if (MODEL) A.torque[i] = tq;
  }
}

// End of a fused force kernel: each wave's reduction of the new state (held
// by every L-th lane) joins a sub-slot of the scenario's slot for the next
// step (no block barrier); the step is counted.
__device__ __forceinline__ void step_end(const Work &W, int scn, int p, Red r, int L) {
  r = wave_reduce(r, L);
  const int wave = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  if ((threadIdx.x & 63) == 0) acc_merge(acc_at(W, scn, 1 - p, wave % kAccSub), r);
  if (blockIdx.x == 0 && threadIdx.x == 0) W.st->step += 1;
}

// Direct kernel: blocks of blockDim / L agents of one scenario, in agent
// index order, L lanes each.  FUSED = false: the AgentAgentInteractions
// node, force (torque) += pairs; nothing runs when the scenario's binning
// is invalid.  FUSED = true: the pair and wall forces of one fused step
// (after the own terms the step's first kernel added), in the post-order of
// the standard logic tree, then the integrator and Reset of each agent.
// Neighbour state comes from the step's snapshot, because agents move as
// they are integrated; without a valid binning only the walls are added.
template <typename T, int MODEL, bool FUSED, bool ONE>
__global__ void __launch_bounds__(256, MODEL == 0 ? 3 : 1) k_forces_direct(Agents<T> A, Work W, Batch B,
                                                       FusedArgs fa, int L, int reset,
                                                       int ordered) {
  __shared__ double4 sseg[3 * kWallTile];
  const int2 sk = block_e_t<ONE>(B);
  const ScnDesc d = desc_t<ONE>(B, sk.x);
  // the walls' staging first: its wave waits for the segments, and its
  // loads would otherwise queue behind the agent's
  stage_walls(fa, d, sseg, FUSED);
  const int per = (int)blockDim.x >> __builtin_ctz(L);  // L is a power of two
  // the agent: the next in index order, or in the processing order (an
  // entry outside the scenario, which a valid order never holds, serves none)
  const int local = sk.y * per + ((int)threadIdx.x >> __builtin_ctz(L));
  int i = d.a0 + local;
  if (ordered && i < d.a1) i = W.order[i];
  const bool have = i >= d.a0 && i < d.a1;
  const bool lead = have && (threadIdx.x & (L - 1)) == 0;
  double2 pos = make_double2(0.0, 0.0);
  V2<T> vel;
  vel.x = T(0);
  vel.y = T(0);
  int c = d.c0;
  int self_slot = -1;
  TCView me;
  Self<T> self;
  Tail<T> tail;
  // The agent's loads come first: they need only its index, not the
  // scenario's state, whose loads would otherwise each wait for the one
  // before.  The cell and slot are read whatever the binning's validity
  // and used only when it holds.
  if (have) {
    pos = A.pos[i];
    vel = A.vel[i];
    c = W.cell_of[i];
    if (MODEL == 0 && FUSED) self_slot = W.rank[i];
    if constexpr (MODEL == 1) me = tc_agent(A, i);
    self = self_of<T, MODEL>(A, i, pos);
    if (FUSED && lead) tail = tail_of(A, i);
  }
  // the scenario's state, every field read unconditionally (a short-circuit
  // test would hold each load until the previous one arrived)
  const ScnState *ss = &B.sst[sk.x];
  const int par = ss->par, valid = ss->valid, oob0 = ss->oob[0], oob1 = ss->oob[1];
  const int nx = ss->nx, ny = ss->ny;
  const double dt_s = ss->dt;
  const int p = FUSED ? par : 0;
  const bool pairs_on = FUSED ? ((fa.flags & CD_USE_AGENT_AGENT) != 0) : true;
  const bool ok = pairs_on & (valid != 0) & ((p ? oob1 : oob0) == 0);
  if (!FUSED && !ok) return;
  const double dt = FUSED ? dt_s : 0.0;
  if (!ok) {
    c = d.c0;
    self_slot = -1;
  }
  const int lc = c - d.c0, lx = ny > 0 ? lc / ny : 0, ly = lc - lx * ny;
  const MemRow<T, MODEL> rows{A, W, W.count + (size_t)p * W.cap,
                              FUSED ? nbr_snapshot<T>(W) : nbr_agents(A),
                              d.c0, nx, ny, lx, ly, d.a0, d.a1,
                              (MODEL == 0 && FUSED) ? (const NbrRec<T> *)W.rec : nullptr,
                              self_slot};
  __syncthreads();  // walls staged
  V2<T> f;
  double tq;
  agent_forces<T, MODEL, FUSED>(fa, d, sseg, L, have, ok, i, pos, vel, me, self, rows, f, tq);
  Red r;
  r.init();
  agent_finish<T, MODEL, FUSED>(A, fa, i, lead, f, tq, pos, vel, tail, dt, reset, r);
  if constexpr (FUSED) step_end(W, sk.x, p, r, L);
}

struct TileArgs {
  int tx, ty;  // cells per tile
};

// What the tiled kernel's agents need beyond the force pass: the count
// buffer of the binning, the neighbour state, and for the fused step dt and
// Reset.
template <typename T>
struct TileCtx {
  const int *count;
  Nbr<T> N;
  double dt;
  int reset;
};

// Stages the region of w x h inner cells at grid cell (cx0, cy0) and serves
// its agents.  Returns false, having done nothing, when the region holds
// more agents than the LDS capacity.
template <typename T, int MODEL, bool FUSED>
__device__ bool tile_region(const Agents<T> &A, const Work &W, const FusedArgs &fa,
                            const ScnDesc &d, int nx, int ny, TileLds<T, MODEL> &S,
                            double4 *sseg, int cx0, int cy0, int w, int h,
                            const TileCtx<T> &cx, Red &red) {
  constexpr int CAP = TileLds<T, MODEL>::CAP;
  const int tid = threadIdx.x;
  const int RH = h + 2, NC = (w + 2) * RH;
  const int rx0 = cx0 - 1, ry0 = cy0 - 1;
  auto gcell = [&](int k) { return d.c0 + (rx0 + k / RH) * ny + (ry0 + k % RH); };
  auto count_of = [&](int k) -> int {
    if (k >= NC) return 0;
    const int lx = rx0 + k / RH, ly = ry0 + k % RH;
    if (lx < 0 || lx >= nx || ly < 0 || ly >= ny) return 0;
    return cx.count[d.c0 + lx * ny + ly];
  };
  // 1. counts and offsets (two cells per thread: NC <= 2 * kTileThreads)
  const int k0 = 2 * tid, k1 = 2 * tid + 1;
  const int n0 = count_of(k0), n1 = count_of(k1);
  const int over = __syncthreads_or(n0 > W.K || n1 > W.K);
  int M;
  const int x0 = block_scan_excl(n0 + n1, S.wsum, M);
  if (M > CAP) return false;
  if (k0 < NC) {
    S.cnt[k0] = n0;
    S.off[k0] = x0;
  }
  if (k1 < NC) {
    S.cnt[k1] = n1;
    S.off[k1] = x0 + n0;
  }
  if (tid == 0) S.off[NC] = M;
  __syncthreads();
  // 2. ids from the slot lists
  for (int e = tid; e < M; e += kTileThreads) {
    int lo = 0, hi = NC - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (S.off[mid] <= e)
        lo = mid;
      else
        hi = mid - 1;
    }
    S.cell[e] = lo;
    if (S.cnt[lo] <= W.K) S.id[e] = W.ell_ids[(size_t)gcell(lo) * W.K + (e - S.off[lo])];
  }
  if (over) {  // members of overfull cells, compacted by an ordered block scan
    for (int k = 0; k < NC; ++k) {
      if (S.cnt[k] <= W.K) continue;
      const int gc = gcell(k);
      int base = S.off[k];
      for (int j0 = d.a0; j0 < d.a1; j0 += kTileThreads) {
        const int j = j0 + tid;
        const int pred = (j < d.a1 && W.cell_of[j] == gc) ? 1 : 0;
        int tot;
        const int r = block_scan_excl(pred, S.wsum, tot);
        if (pred) S.id[base + r] = j;
        base += tot;
      }
    }
  }
  __syncthreads();
  // 3. neighbour data
  for (int e = tid; e < M; e += kTileThreads) {
    const int j = S.id[e];
    S.pos[e] = cx.N.pos[j];
    S.vel[e] = cx.N.vel[j];
    S.r[e] = A.radius[j];
    if constexpr (MODEL == 1) {
      S.ls[e] = cx.N.ls[j];
      S.rs[e] = cx.N.rs[j];
      S.rt[e] = A.r_t[j];
      S.rsr[e] = A.r_s[j];
      S.sho[e] = cx.N.off ? cx.N.off[j] : shoulder_offset(A.r_ts[j], cx.N.phi[j]);
    }
  }
  if (tid == 0) {  // own agents: the inner cells, column by column
    int acc = 0;
    for (int c = 0; c < w; ++c) {
      S.own[c] = acc;
      acc += S.off[(c + 1) * RH + h + 1] - S.off[(c + 1) * RH + 1];
    }
    S.own[w] = acc;
  }
  __syncthreads();
  // 4. the agents of the inner cells, four lanes each
  const int Mo = S.own[w];
  for (int base = 0; base < Mo; base += kTileGroups) {
    const int o = base + (tid >> 2);
    const bool have = o < Mo;
    int e = 0;
    if (have) {
      int c = 0;
      while (c + 1 < w && S.own[c + 1] <= o) ++c;
      e = S.off[(c + 1) * RH + 1] + (o - S.own[c]);
    }
    const int k = S.cell[e];
    const int kx = k / RH;
    const LdsRow<T, MODEL> rows{S, RH, kx, k - kx * RH};
    TCView me;
    if constexpr (MODEL == 1) me = tc_lds<T, MODEL>(S, e);
    const int i = S.id[e];
    const double2 pos = S.pos[e];
    const V2<T> vel = S.vel[e];
    const bool lead = have && (tid & 3) == 0;
    Self<T> self;
    Tail<T> tail;
    if (have) self = self_of<T, MODEL>(A, i, pos);
    if (FUSED && lead) tail = tail_of(A, i);
    V2<T> f;
    double tq;
    agent_forces<T, MODEL, FUSED>(fa, d, sseg, 4, have, true, i, pos, vel, me, self, rows, f, tq);
    agent_finish<T, MODEL, FUSED>(A, fa, i, lead, f, tq, pos, vel, tail, cx.dt, cx.reset, red);
  }
  __syncthreads();
  return true;
}

// Serves the agents of grid cell (lx, ly) from memory (exact slow path).
// Also the fused step's fallback without a valid binning (pairs = false,
// gc ignored): every agent of the scenario from block part `part` of `nb`,
// walls and integration only.
template <typename T, int MODEL, bool FUSED>
__device__ void cell_from_memory(const Agents<T> &A, const Work &W, const FusedArgs &fa,
                                 const ScnDesc &d, int nx, int ny, double4 *sseg, int lx,
                                 int ly, const TileCtx<T> &cx, Red &red, bool pairs = true,
                                 int part = 0, int nb = 1) {
  const int gc = d.c0 + lx * ny + ly;
  const MemRow<T, MODEL> rows{A, W, cx.count, cx.N, d.c0, nx, ny, lx, ly, d.a0, d.a1, nullptr, -1};
  for (int j0 = d.a0 + part * kTileGroups; j0 < d.a1; j0 += nb * kTileGroups) {
    const int i = j0 + (threadIdx.x >> 2);
    const bool have = i < d.a1 && (!pairs || W.cell_of[i] == gc);
    const bool lead = have && (threadIdx.x & 3) == 0;
    double2 pos = make_double2(0.0, 0.0);
    V2<T> vel;
    vel.x = T(0);
    vel.y = T(0);
    TCView me;
    Self<T> self;
    Tail<T> tail;
    if (have) {
      pos = A.pos[i];
      vel = A.vel[i];
      if constexpr (MODEL == 1) me = tc_agent(A, i);
      self = self_of<T, MODEL>(A, i, pos);
      if (FUSED && lead) tail = tail_of(A, i);
    }
    V2<T> f;
    double tq;
    agent_forces<T, MODEL, FUSED>(fa, d, sseg, 4, have, pairs, i, pos, vel, me, self, rows, f,
                                  tq);
    agent_finish<T, MODEL, FUSED>(A, fa, i, lead, f, tq, pos, vel, tail, cx.dt, cx.reset, red);
  }
}

// Tiled kernel (same contract as k_forces_direct): blocks of a scenario
// stride over its tiles; without a valid binning (fused step only) they
// stride over its agents.
template <typename T, int MODEL, bool FUSED>
__global__ void __launch_bounds__(kTileThreads) k_forces_tiled(Agents<T> A, Work W, Batch B,
                                                                FusedArgs fa, TileArgs ta,
                                                                int reset) {
  __shared__ TileLds<T, MODEL> S;
  __shared__ double4 sseg[3 * kWallTile];
  const int2 sk = B.map_t ? B.map_t[blockIdx.x] : make_int2(0, (int)blockIdx.x);
  const ScnDesc d = desc_of(B, sk.x);
  const int nb = B.map_t ? d.nt : (int)gridDim.x;
  ScnState *ss = &B.sst[sk.x];
  const int p = FUSED ? ss->par : 0;
  const bool pairs_on = FUSED ? ((fa.flags & CD_USE_AGENT_AGENT) != 0) : true;
  const bool ok = pairs_on && ss->valid && !ss->oob[p];
  if (!FUSED && !ok) return;
  TileCtx<T> cx;
  cx.count = W.count + (size_t)p * W.cap;
  cx.N = FUSED ? nbr_snapshot<T>(W) : nbr_agents(A);
  cx.dt = FUSED ? ss->dt : 0.0;
  cx.reset = reset;
  stage_walls(fa, d, sseg, FUSED);
  __syncthreads();
  Red red;
  red.init();
  if (ok) {
    const int nx = ss->nx, ny = ss->ny;
    const int ntx = (nx + ta.tx - 1) / ta.tx, nty = (ny + ta.ty - 1) / ta.ty;
    const int ntiles = ntx * nty;
    for (int t = sk.y; t < ntiles; t += nb) {
      const int cx0 = (t / nty) * ta.tx, cy0 = (t % nty) * ta.ty;
      const int w = min(ta.tx, nx - cx0), h = min(ta.ty, ny - cy0);
      if (tile_region<T, MODEL, FUSED>(A, W, fa, d, nx, ny, S, sseg, cx0, cy0, w, h, cx, red))
        continue;
      for (int c = 0; c < w * h; ++c) {
        const int lx = cx0 + c / h, ly = cy0 + c % h;
        if (!tile_region<T, MODEL, FUSED>(A, W, fa, d, nx, ny, S, sseg, lx, ly, 1, 1, cx, red))
          cell_from_memory<T, MODEL, FUSED>(A, W, fa, d, nx, ny, sseg, lx, ly, cx, red);
      }
    }
  } else {
    cell_from_memory<T, MODEL, FUSED>(A, W, fa, d, 0, 0, sseg, 0, 0, cx, red, false, sk.y, nb);
  }
  if constexpr (FUSED) step_end(W, sk.x, p, red, 4);
}

// ---------------------------------------------------------------------------
// Integrator node
// ---------------------------------------------------------------------------

// The integrator of agent i with its force (and torque) given, see
// verlet_with.
template <typename T>
__device__ __forceinline__ void verlet_agent(const Agents<T> &A, int model, int i, double dt,
                                             bool do_reset, double2 &x_out, V2<T> &v_out) {
  verlet_with(A, model, i, dt, do_reset, A.force[i], model ? A.torque[i] : 0.0, A.pos[i],
              A.vel[i], A.force_prev[i], A.mass[i], x_out, v_out);
}

// Integrator node: dt from the partials the preceding k_reduce wrote.
template <typename T>
__global__ void __launch_bounds__(256) k_integrate_node(Agents<T> A, int model, Work W, Batch B) {
  const int2 sk = block_f(B);
  const ScnDesc d = desc_of(B, sk.x);
  ScnState *ss = &B.sst[sk.x];
  const Red r = reduce_partials(W, d);
  const double dt = adaptive_dt(r.v, d.dt_min, d.dt_max);
  if (sk.y == 0 && threadIdx.x == 0) {
    for (int k = 0; k < R_COUNT; ++k) ss->red[k] = r.v[k];
    ss->dt = dt;
    ss->time_tot += dt;
  }
  const int stride = d.nf * blockDim.x;
  for (int i = d.a0 + sk.y * blockDim.x + threadIdx.x; i < d.a1; i += stride) {
    double2 x;
    V2<T> v;
    verlet_agent(A, model, i, dt, false, x, v);
  }
}

// ---------------------------------------------------------------------------
// Host helpers
// ---------------------------------------------------------------------------


template <typename T, int MODEL, bool FUSED>
void launch_forces_m(const Agents<T> &A, const Work &W, const Batch &B, const FusedArgs &fa,
                     const Launch &L, const cd_workspace_t *w, hipStream_t s, int reset) {
  if (w->tile_x > 0 && w->tile_y > 0) {
    TileArgs ta;
    ta.tx = (int)w->tile_x;
    ta.ty = (int)w->tile_y;
    hipLaunchKernelGGL((k_forces_tiled<T, MODEL, FUSED>), dim3(L.blocks_t), dim3(kTileThreads), 0,
                       s, A, W, B, fa, ta, reset);
  } else if (!B.desc || !B.map_e) {  // one scenario (single_batch: table and maps null)
    hipLaunchKernelGGL((k_forces_direct<T, MODEL, FUSED, true>), dim3(L.blocks_e),
                       dim3(L.threads_e), 0, s, A, W, B, fa, L.lanes, reset,
                       (FUSED && W.order_every > 0) ? 1 : 0);
  } else {
    hipLaunchKernelGGL((k_forces_direct<T, MODEL, FUSED, false>), dim3(L.blocks_e),
                       dim3(L.threads_e), 0, s, A, W, B, fa, L.lanes, reset,
                       (FUSED && W.order_every > 0) ? 1 : 0);
  }
}

// The force kernel: of the fused step (the step's second and last kernel,
// forces then integration; `reset` as CD_USE_RESET), or of the
// AgentAgentInteractions node.
template <typename T>
int launch_forces(const cd_agents_t *a, const Work &W, const Batch &B, const FusedArgs &fa,
                  bool fused, const Launch &L, const cd_workspace_t *w, hipStream_t s,
                  int reset = 0) {
  const Agents<T> A = typed<T>(a);
  if (a->model == CD_THREE_CIRCLE) {
    if (fused)
      launch_forces_m<T, 1, true>(A, W, B, fa, L, w, s, reset);
    else
      launch_forces_m<T, 1, false>(A, W, B, fa, L, w, s, reset);
  } else {
    if (fused)
      launch_forces_m<T, 0, true>(A, W, B, fa, L, w, s, reset);
    else
      launch_forces_m<T, 0, false>(A, W, B, fa, L, w, s, reset);
  }
  return check_launch();
}

Launch batch_launch(const cd_workspace_t *w) {
  Launch L = single_launch(w);
  L.blocks_f = w->blocks_f;
  L.blocks_e = std::max(1, (int)w->blocks_e);
  L.blocks_t = std::max(1, (int)w->blocks_t);
  L.scenarios = std::max(1, (int)w->n_scenarios);
  return L;
}


// K_a of the fused step, for one scenario (ONE) or a batch.
template <typename T, bool ONE>
void launch_step_begin_m(const Agents<T> &A, const cd_step_params_t *p, const Work &W,
                         const Batch &B, const FusedArgs &fa, const Launch &L, hipStream_t s,
                         int model, int binning, bool split, dim3 grid_s, dim3 block_s) {
  if (model == CD_THREE_CIRCLE) {
    if (split)
      hipLaunchKernelGGL((k_step_begin<T, 1, true, ONE>), grid_s, block_s, 0, s, A, W, B, fa,
                         p->cell_size, binning);
    else
      hipLaunchKernelGGL((k_step_begin<T, 1, false, ONE>), dim3(L.blocks_f), dim3(kBS), 0, s, A,
                         W, B, fa, p->cell_size, binning);
  } else if (split) {
    hipLaunchKernelGGL((k_step_begin<T, 0, true, ONE>), grid_s, block_s, 0, s, A, W, B, fa,
                       p->cell_size, binning);
  } else if (B.map_f) {
    hipLaunchKernelGGL((k_step_begin<T, 0, false, ONE>), dim3(L.blocks_f), dim3(kBS), 0, s, A, W,
                       B, fa, p->cell_size, binning);
  } else {  // one scenario: one-wave blocks, so the waves spread over all CUs
    hipLaunchKernelGGL((k_step_begin<T, 0, false, ONE>), dim3(4 * L.blocks_f), dim3(64), 0, s, A,
                       W, B, fa, p->cell_size, binning);
  }
}

// K_a of the fused step.
template <typename T>
void launch_step_begin(const cd_agents_t *a, const cd_step_params_t *p, const Work &W,
                       const Batch &B, const FusedArgs &fa, const Launch &L, hipStream_t s) {
  const Agents<T> A = typed<T>(a);
  const int binning = (p->flags & CD_USE_AGENT_AGENT) ? 1 : 0;
  // with fluctuation or navigation: a binning and an own-terms half per
  // block over the same agents (one scenario: 128-thread blocks, 64 agents
  // each, so the waves spread over all CUs; a batch: the block map's
  // 256-thread blocks)
  const bool split = (p->flags & (CD_USE_FLUCT | CD_USE_NAV)) != 0;
  const bool one = !B.map_f;
  const dim3 grid_s(one ? 4 * L.blocks_f : L.blocks_f), block_s(one ? 128 : kBS);
  if (one || !B.desc)  // single_batch: table and maps null
    launch_step_begin_m<T, true>(A, p, W, B, fa, L, s, (int)a->model, binning, split, grid_s,
                                 block_s);
  else
    launch_step_begin_m<T, false>(A, p, W, B, fa, L, s, (int)a->model, binning, split, grid_s,
                                  block_s);
  if (W.order_every > 0 && binning)
    hipLaunchKernelGGL(k_order_build, dim3(L.scenarios), dim3(kOrderThreads), 0, s, W, B);
}

// One step: K_a, then the force kernel (forces and integration); events
// ev[0..3] around them when timed (ev[2] and ev[3] both after the force
// kernel: the step has no separate integration kernel).
template <typename T>
int step_kernels(const cd_agents_t *a, const cd_step_params_t *p, const Work &W, const Batch &B,
                 const Launch &L, const cd_workspace_t *w, hipStream_t s, hipEvent_t *ev,
                 double *tape_x = nullptr, double *tape_v = nullptr) {
  FusedArgs fa = fused_args(p);
  fa.tape_x = (double2 *)tape_x;
  fa.tape_v = (void *)tape_v;
  const int reset = (p->flags & CD_USE_RESET) ? 1 : 0;
  mark(ev, 0, s);
  launch_step_begin<T>(a, p, W, B, fa, L, s);
  mark(ev, 1, s);
  const int rc = launch_forces<T>(a, W, B, fa, true, L, w, s, reset);
  if (rc) return rc;
  mark(ev, 2, s);
  mark(ev, 3, s);
  return check_launch();
}

template <typename T>
int do_agent_agent(const cd_agents_t *a, double cell, const cd_workspace_t *w, hipStream_t s) {
  const Agents<T> A = typed<T>(a);
  const Work W = carve(w);
  const Batch B = single_batch(w, 0, 0, 0.0, 0.0);
  const Launch L = single_launch(w);
  int rc;
  if ((rc = launch_reduce(A, W, B, L.blocks_f, true, s))) return rc;
  hipLaunchKernelGGL(k_bin_insert, dim3(L.blocks_f), dim3(kBS), 0, s, (const double2 *)A.pos,
                     cell, W, B, 1);
  if ((rc = check_launch())) return rc;
  FusedArgs fa;
  memset(&fa, 0, sizeof(fa));
  return launch_forces<T>(a, W, B, fa, false, L, w, s);
}

template <typename T>
int do_add_to_cells(const cd_agents_t *a, double cell, const cd_workspace_t *w, hipStream_t s) {
  const Agents<T> A = typed<T>(a);
  const Work W = carve(w);
  const Batch B = single_batch(w, 0, 0, 0.0, 0.0);
  const Launch L = single_launch(w);
  int rc;
  if ((rc = launch_reduce(A, W, B, L.blocks_f, true, s))) return rc;
  hipLaunchKernelGGL(k_bl_count, dim3(L.blocks_f), dim3(kBS), 0, s, (const double2 *)A.pos,
                     cell, W, B);
  hipLaunchKernelGGL(k_bl_scan, dim3(W.ntiles), dim3(kScanThreads), 0, s, W, B.sst);
  hipLaunchKernelGGL(k_bl_scatter, dim3(grid_for(A.n, kBS)), dim3(kBS), 0, s, A.n, W,
                     (const ScnState *)B.sst);
  hipLaunchKernelGGL(k_bl_sort, dim3(grid_for(W.cap, kBS)), dim3(kBS), 0, s, W,
                     (const ScnState *)B.sst);
  return check_launch();
}


// Scenario table and block maps of a batch (host side).
struct HostBatch {
  std::vector<ScnDesc> desc;
  std::vector<int2> map_f, map_e, map_t;
};

int build_batch(const cd_scenario_t *sc, int S, const cd_workspace_t *w, int64_t n_segments,
                HostBatch &H) {
  H.desc.resize(S);
  int f = 0;
  int64_t prev_end = 0;
  for (int s = 0; s < S; ++s) {
    const cd_scenario_t &c = sc[s];
    if (c.agent_begin < prev_end || c.agent_end < c.agent_begin || c.agent_end > w->n ||
        c.cell_begin < 0 || c.cell_capacity <= 0 ||
        c.cell_begin + c.cell_capacity > w->cell_capacity || c.seg_begin < 0 ||
        c.seg_end < c.seg_begin || c.seg_end > n_segments || !(c.dt_min >= 0.0) ||
        !(c.dt_max >= 0.0))
      return CD_ERR_ARG;
    prev_end = c.agent_end;
    ScnDesc &d = H.desc[s];
    d.a0 = (int)c.agent_begin;
    d.a1 = (int)c.agent_end;
    d.c0 = (int)c.cell_begin;
    d.ccap = (int)c.cell_capacity;
    d.w0 = (int)c.seg_begin;
    d.w1 = (int)c.seg_end;
    d.seed = c.seed;
    d.dt_min = c.dt_min;
    d.dt_max = c.dt_max;
    const int64_t ns = c.agent_end - c.agent_begin;
    d.nf = std::max(1, std::min(grid_for(ns, kBS), kMaxPartsPerScn));
    d.f0 = f;
    for (int k = 0; k < d.nf; ++k) H.map_f.push_back(make_int2(s, k));
    f += d.nf;
    const int ne = grid_for(ns, force_threads(w) / agent_lanes(w));
    for (int k = 0; k < ne; ++k) H.map_e.push_back(make_int2(s, k));
    d.nt = tile_blocks(c.cell_capacity, w);
    for (int k = 0; k < d.nt; ++k) H.map_t.push_back(make_int2(s, k));
  }
  const int64_t Smax = std::max<int64_t>(1, w->max_scenarios);
  if ((int64_t)H.map_f.size() > max_blocks_f(w->n, Smax) ||
      (int64_t)H.map_e.size() > max_blocks_e(w->n, Smax) ||
      (int64_t)H.map_t.size() > w->cell_capacity + Smax + 1)
    return CD_ERR_WORKSPACE;
  return CD_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free AgentAgentInteractions: the workspace of
// cd_agent_agent_block_list lives in `scratch`, a device array the library
// holds on every GPU it runs on until the process ends.  Host code reaches
// the array only through the address hipGetSymbolAddress returns for the
// current device (kept per device below); the array's own name is not a
// device address on the host.
// ---------------------------------------------------------------------------

constexpr int64_t kScratchBytes = CD_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of `scratch` (looked up once per device).
int scratch_ptr(void **out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return CD_ERR_HIP;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return CD_ERR_HIP;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return CD_OK;
}

// The call's status bits to the caller's word (0 when st is null: no agents).
__global__ void k_status_out(const DevState *st, int64_t *out) {
  *out = st ? (int64_t)st->status : 0;
}

__global__ void k_clear_status(DevState *st) { st->status = 0; }
__global__ void k_set_step(DevState *st, unsigned long long step) { st->step = step; }
__global__ void __launch_bounds__(256) k_hypot(const double2 *xy, double *out, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = hypot_ref(xy[i].x, xy[i].y);
}
__global__ void k_philox(const uint32_t *in, uint32_t *out) {
  uint32_t c[4] = {in[0], in[1], in[2], in[3]};
  philox4x32(c, in[4], in[5]);
  for (int k = 0; k < 4; ++k) out[k] = c[k];
}

void to_state(const DevState &g, const ScnState &h, cd_state_t *out) {
  out->dt = h.dt;
  out->time_tot = h.time_tot;
  out->origin[0] = h.origin[0];
  out->origin[1] = h.origin[1];
  out->shape[0] = h.nx;
  out->shape[1] = h.ny;
  out->status = g.status;
  out->step = (int64_t)g.step;
  double vm;
  memcpy(&vm, &h.red[R_VMAX], 8);
  out->v_max = vm;
  unsigned long long b = h.red[R_V0MAX];
  b = (b & 0x8000000000000000ull) ? (b & 0x7fffffffffffffffull) : ~b;
  memcpy(&vm, &b, 8);
  out->v0_max = vm;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

int64_t cd_workspace_bytes(const cd_workspace_t *w, cd_layout_t *layout) {
  if (!ws_sizes_ok(w)) return -1;
  const Layout L = make_layout(w);
  if (layout) {
    layout->points_indices = L.pidx;
    layout->cells_count = L.count;
    layout->cells_offset = L.offset;
    layout->cell_of = L.cell_of;
    layout->total_bytes = L.total;
  }
  return L.total;
}

int cd_workspace_init(const cd_workspace_t *w, cd_stream_t stream) {
  if (!ws_sizes_ok(w) || !w->ptr) return CD_ERR_ARG;
  const Layout L = make_layout(w);
  if (w->bytes < L.total) return CD_ERR_WORKSPACE;
  const Work W = carve(w);
  const int S = (int)std::max<int64_t>(1, w->max_scenarios);
  hipLaunchKernelGGL(k_ws_init, dim3(std::min(grid_for(w->cell_capacity + S + 1, kBS), 4096)),
                     dim3(kBS), 0, (hipStream_t)stream, W, (ScnState *)((char *)w->ptr + L.scn_state),
                     S);
  return check_launch();
}

int cd_batch_read_states(const cd_workspace_t *w, cd_state_t *out, int32_t count,
                         cd_stream_t stream) {
  if (!w || !w->ptr || !out || count < 1 || count > std::max<int64_t>(1, w->max_scenarios))
    return CD_ERR_ARG;
  const Layout L = make_layout(w);
  DevState g;
  std::vector<ScnState> h(count);
  hipStream_t s = (hipStream_t)stream;
  if (hipMemcpyAsync(&g, w->ptr, sizeof(DevState), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(h.data(), (char *)w->ptr + L.scn_state, sizeof(ScnState) * count,
                     hipMemcpyDeviceToHost, s) != hipSuccess)
    return CD_ERR_HIP;
  if (hipStreamSynchronize(s) != hipSuccess) return CD_ERR_HIP;
  for (int k = 0; k < count; ++k) to_state(g, h[k], &out[k]);
  return CD_OK;
}

int cd_read_state(const cd_workspace_t *w, cd_state_t *out, cd_stream_t stream) {
  if (!ws_sizes_ok(w)) return CD_ERR_ARG;
  return cd_batch_read_states(w, out, 1, stream);
}

int cd_clear_status(const cd_workspace_t *w, cd_stream_t stream) {
  if (!w || !w->ptr) return CD_ERR_ARG;
  hipLaunchKernelGGL(k_clear_status, dim3(1), dim3(1), 0, (hipStream_t)stream,
                     (DevState *)w->ptr);
  return check_launch();
}

int cd_add_to_cells(const cd_agents_t *a, double cell_size, const cd_workspace_t *w,
                    cd_stream_t stream) {
  int rc = check_agents(a);
  if (rc) return rc;
  if (!(cell_size > 0.0)) return CD_ERR_ARG;
  if ((rc = check_ws(a, w))) return rc;
  if (a->n == 0) return CD_OK;
  if (a->real == CD_F64) return do_add_to_cells<double>(a, cell_size, w, (hipStream_t)stream);
  return do_add_to_cells<float>(a, cell_size, w, (hipStream_t)stream);
}

int cd_agent_agent_block_list(const cd_agents_t *a, double cell_size, const cd_workspace_t *w,
                              cd_stream_t stream) {
  int rc = check_agents(a);
  if (rc) return rc;
  if (!(cell_size > 0.0)) return CD_ERR_ARG;
  if ((rc = check_ws(a, w))) return rc;
  if (a->n == 0) return CD_OK;
  if (a->real == CD_F64) return do_agent_agent<double>(a, cell_size, w, (hipStream_t)stream);
  return do_agent_agent<float>(a, cell_size, w, (hipStream_t)stream);
}

int cd_agent_agent_block_list_noalloc(const cd_agents_t *a, double cell_size,
                                      int64_t cell_capacity, int64_t cell_slots,
                                      int64_t *status_out, cd_stream_t stream) {
  int rc = check_agents(a);
  if (rc) return rc;
  if (!(cell_size > 0.0)) return CD_ERR_ARG;
  cd_workspace_t w;
  memset(&w, 0, sizeof(w));
  w.n = a->n;
  w.cell_capacity = cell_capacity;
  w.cell_slots = cell_slots;
  w.model = a->model;
  if (!ws_sizes_ok(&w)) return CD_ERR_ARG;
  if (make_layout(&w).total > kScratchBytes) return CD_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  if (a->n == 0) {
    if (status_out) {
      hipLaunchKernelGGL(k_status_out, dim3(1), dim3(1), 0, s, (const DevState *)nullptr,
                         status_out);
      return check_launch();
    }
    return CD_OK;
  }
  void *p = nullptr;
  if ((rc = scratch_ptr(&p))) return rc;
  w.ptr = p;
  w.bytes = kScratchBytes;
  // the previous call on this GPU left its state, counts and scan tiles
  // (at the offsets of its own sizes): clear them for this call's layout
  const Work W = carve(&w);
  const Layout L = make_layout(&w);
  hipLaunchKernelGGL(k_ws_init, dim3(std::min(grid_for(w.cell_capacity + 2, kBS), 4096)),
                     dim3(kBS), 0, s, W, (ScnState *)((char *)p + L.scn_state), 1);
  if ((rc = check_launch())) return rc;
  rc = (a->real == CD_F64) ? do_agent_agent<double>(a, cell_size, &w, s)
                           : do_agent_agent<float>(a, cell_size, &w, s);
  if (rc) return rc;
  if (status_out) {
    hipLaunchKernelGGL(k_status_out, dim3(1), dim3(1), 0, s, (const DevState *)W.st,
                       status_out);
    return check_launch();
  }
  return CD_OK;
}

int cd_agent_obstacle(const cd_agents_t *a, const double *segments, int64_t W,
                      cd_stream_t stream) {
  int rc = check_agents(a);
  if (rc) return rc;
  if (W < 0 || (W > 0 && !segments) || W > 0x7fffffff) return CD_ERR_ARG;
  if (a->n == 0 || W == 0) return CD_OK;
  hipStream_t s = (hipStream_t)stream;
  if (a->real == CD_F64)
    hipLaunchKernelGGL(k_walls<double>, dim3(grid_for(a->n, kBS)), dim3(kBS), 0, s,
                       typed<double>(a), (int)a->model, (const double4 *)segments, (int)W);
  else
    hipLaunchKernelGGL(k_walls<float>, dim3(grid_for(a->n, kBS)), dim3(kBS), 0, s,
                       typed<float>(a), (int)a->model, (const double4 *)segments, (int)W);
  return check_launch();
}

int cd_force_adjust_agents(const cd_agents_t *a, cd_stream_t stream) {
  int rc = check_agents(a);
  if (rc || a->n == 0) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (a->real == CD_F64)
    hipLaunchKernelGGL(k_adjust<double>, dim3(grid_for(a->n, kBS)), dim3(kBS), 0, s,
                       typed<double>(a), 1, 0);
  else
    hipLaunchKernelGGL(k_adjust<float>, dim3(grid_for(a->n, kBS)), dim3(kBS), 0, s,
                       typed<float>(a), 1, 0);
  return check_launch();
}

int cd_torque_adjust_agents(const cd_agents_t *a, cd_stream_t stream) {
  int rc = check_agents(a);
  if (rc || a->n == 0) return rc;
  if (a->model != CD_THREE_CIRCLE) return CD_ERR_TYPE;
  hipLaunchKernelGGL(k_adjust<double>, dim3(grid_for(a->n, kBS)), dim3(kBS), 0,
                     (hipStream_t)stream, typed<double>(a), 0, 1);
  return check_launch();
}

int cd_orient_towards_target_direction(const cd_agents_t *a, cd_stream_t stream) {
  int rc = check_agents(a);
  if (rc || a->n == 0) return rc;
  if (a->model != CD_THREE_CIRCLE) return CD_ERR_TYPE;
  hipLaunchKernelGGL(k_orient<double>, dim3(grid_for(a->n, kBS)), dim3(kBS), 0,
                     (hipStream_t)stream, typed<double>(a));
  return check_launch();
}

int cd_navigation(const cd_agents_t *a, const double *maps, int64_t T_, int64_t rows,
                  int64_t cols, double minx, double miny, double step, int32_t mode,
                  cd_stream_t stream) {
  int rc = check_agents(a);
  if (rc) return rc;
  if (!maps || T_ < 0 || rows < 1 || cols < 1 || !(step > 0.0) || (mode != 0 && mode != 1))
    return CD_ERR_ARG;
  if (mode == 1 && (rows < 2 || cols < 2)) return CD_ERR_ARG;
  if (a->n == 0 || T_ == 0) return CD_OK;
  hipStream_t s = (hipStream_t)stream;
  if (a->real == CD_F64)
    hipLaunchKernelGGL(k_nav<double>, dim3(grid_for(a->n, kBS)), dim3(kBS), 0, s,
                       typed<double>(a), maps, T_, rows, cols, minx, miny, step, (int)mode);
  else
    hipLaunchKernelGGL(k_nav<float>, dim3(grid_for(a->n, kBS)), dim3(kBS), 0, s,
                       typed<float>(a), maps, T_, rows, cols, minx, miny, step, (int)mode);
  return check_launch();
}

int cd_force_fluctuation(const cd_agents_t *a, uint64_t seed, uint64_t step, cd_stream_t stream) {
  int rc = check_agents(a);
  if (rc || a->n == 0) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (a->real == CD_F64)
    hipLaunchKernelGGL(k_fluct<double>, dim3(grid_for(a->n, kBS)), dim3(kBS), 0, s,
                       typed<double>(a), (int)a->model, seed, step);
  else
    hipLaunchKernelGGL(k_fluct<float>, dim3(grid_for(a->n, kBS)), dim3(kBS), 0, s,
                       typed<float>(a), (int)a->model, seed, step);
  return check_launch();
}

int cd_velocity_verlet_integrator(const cd_agents_t *a, double dt_min, double dt_max,
                                  const cd_workspace_t *w, cd_stream_t stream) {
  int rc = check_agents(a);
  if (rc) return rc;
  if ((rc = check_ws(a, w))) return rc;
  if (!(dt_min >= 0.0) || !(dt_max >= 0.0)) return CD_ERR_ARG;
  if (a->n == 0) return CD_OK;
  hipStream_t s = (hipStream_t)stream;
  const Work W = carve(w);
  const Batch B = single_batch(w, 0, 0, dt_min, dt_max);
  const Launch L = single_launch(w);
  if (a->real == CD_F64) {
    const Agents<double> A = typed<double>(a);
    if ((rc = launch_reduce(A, W, B, L.blocks_f, false, s))) return rc;
    hipLaunchKernelGGL(k_integrate_node<double>, dim3(L.blocks_f), dim3(kBS), 0, s, A,
                       (int)a->model, W, B);
  } else {
    const Agents<float> A = typed<float>(a);
    if ((rc = launch_reduce(A, W, B, L.blocks_f, false, s))) return rc;
    hipLaunchKernelGGL(k_integrate_node<float>, dim3(L.blocks_f), dim3(kBS), 0, s, A,
                       (int)a->model, W, B);
  }
  return check_launch();
}

int cd_reset(const cd_agents_t *a, cd_stream_t stream) {
  int rc = check_agents(a);
  if (rc || a->n == 0) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (a->real == CD_F64)
    hipLaunchKernelGGL(k_reset<double>, dim3(grid_for(a->n, kBS)), dim3(kBS), 0, s,
                       typed<double>(a), (int)a->model);
  else
    hipLaunchKernelGGL(k_reset<float>, dim3(grid_for(a->n, kBS)), dim3(kBS), 0, s,
                       typed<float>(a), (int)a->model);
  return check_launch();
}

int64_t cd_wall_grid_build(const double *segs, int64_t W, double reach, double cell,
                           int32_t *out, int64_t capacity, cd_step_params_t *p) {
  if (W <= 0 || W > 0x7fffffff || !segs || !(reach >= 0.0) || !(cell > 0.0) || !p ||
      capacity < 0)
    return CD_ERR_ARG;
  double bx0 = HUGE_VAL, by0 = HUGE_VAL, bx1 = -HUGE_VAL, by1 = -HUGE_VAL;
  for (int64_t w = 0; w < W; ++w) {
    const double *s = segs + 4 * w;
    for (int k = 0; k < 4; ++k)
      if (!std::isfinite(s[k])) return CD_ERR_ARG;
    bx0 = std::min(bx0, std::min(s[0], s[2]));
    by0 = std::min(by0, std::min(s[1], s[3]));
    bx1 = std::max(bx1, std::max(s[0], s[2]));
    by1 = std::max(by1, std::max(s[1], s[3]));
  }
  // the force kernel skips a segment whose box lies farther than the
  // agent's reach + 1e-6 from its centre; a further 1e-6 covers rounding
  const double g = reach + 2e-6;
  const double minx = bx0 - g, miny = by0 - g;
  const double sx = (bx1 + g - minx) / cell, sy = (by1 + g - miny) / cell;
  if (!(sx < 1e8 && sy < 1e8)) return CD_ERR_ARG;
  const int64_t nx = (int64_t)sx + 1, ny = (int64_t)sy + 1;
  if (nx * ny > ((int64_t)1 << 26)) return CD_ERR_ARG;
  auto range = [&](double lo, double hi, double o, int64_t m, int64_t &a, int64_t &b) {
    a = std::max<int64_t>(0, std::min<int64_t>(m - 1, (int64_t)std::floor((lo - o) / cell)));
    b = std::max<int64_t>(0, std::min<int64_t>(m - 1, (int64_t)std::floor((hi - o) / cell)));
  };
  std::vector<int64_t> off(nx * ny + 1, 0);
  for (int pass = 0; pass < 2; ++pass) {
    std::vector<int64_t> fill;
    if (pass == 1) {
      for (int64_t c = 0; c < nx * ny; ++c) off[c + 1] += off[c];
      const int64_t total = nx * ny + 1 + off[nx * ny];
      if (total > 0x7fffffff) return CD_ERR_ARG;
      p->wall_nx = nx;
      p->wall_ny = ny;
      p->wall_minx = minx;
      p->wall_miny = miny;
      p->wall_cell = cell;
      if (!out || capacity < total) return total;
      for (int64_t c = 0; c <= nx * ny; ++c) out[c] = (int32_t)off[c];
      fill.assign(off.begin(), off.end() - 1);
    }
    for (int64_t w = 0; w < W; ++w) {  // cells in segment order: ids ascend
      const double *s = segs + 4 * w;
      int64_t ix0, ix1, iy0, iy1;
      range(std::min(s[0], s[2]) - g, std::max(s[0], s[2]) + g, minx, nx, ix0, ix1);
      range(std::min(s[1], s[3]) - g, std::max(s[1], s[3]) + g, miny, ny, iy0, iy1);
      for (int64_t ix = ix0; ix <= ix1; ++ix)
        for (int64_t iy = iy0; iy <= iy1; ++iy) {
          const int64_t c = ix * ny + iy;
          if (pass == 0)
            off[c + 1] += 1;
          else
            out[nx * ny + 1 + fill[c]++] = (int32_t)w;
        }
    }
  }
  return nx * ny + 1 + off[nx * ny];
}

int cd_step_prepare(const cd_agents_t *a, const cd_step_params_t *p, const cd_workspace_t *w,
                    uint64_t first_step, cd_stream_t stream) {
  int rc = check_step(a, p, w);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const Work W = carve(w);
  hipLaunchKernelGGL(k_set_step, dim3(1), dim3(1), 0, s, W.st, (unsigned long long)first_step);
  if ((rc = check_launch())) return rc;
  if (a->n == 0) return CD_OK;
  hipLaunchKernelGGL(k_order_identity, dim3(std::min(grid_for(a->n, kBS), 1024)), dim3(kBS), 0, s,
                     W.order, (int)a->n);
  const Batch B = single_batch(w, p->n_segments, p->seed, p->dt_min, p->dt_max);
  const Launch L = single_launch(w);
  const int slot = (int)(first_step & 1ull);
  hipLaunchKernelGGL(k_scn_prepare, dim3(1), dim3(kBS), 0, s, W, B.sst, 1, slot);
  if (a->real == CD_F64) return launch_reduce(typed<double>(a), W, B, L.blocks_f, true, s, slot);
  return launch_reduce(typed<float>(a), W, B, L.blocks_f, true, s, slot);
}

int cd_step_tape(const cd_agents_t *a, const cd_step_params_t *p, const cd_workspace_t *w,
                 double *tape_x, void *tape_v, cd_stream_t stream) {
  int rc = check_step(a, p, w);
  if (rc) return rc;
  if (a->n == 0) return CD_OK;
  const Work W = carve(w);
  const Batch B = single_batch(w, (p->flags & CD_USE_OBSTACLE) ? p->n_segments : 0, p->seed,
                               p->dt_min, p->dt_max);
  const Launch L = single_launch(w);
  if (a->real == CD_F64)
    return step_kernels<double>(a, p, W, B, L, w, (hipStream_t)stream, nullptr, tape_x,
                                (double *)tape_v);
  return step_kernels<float>(a, p, W, B, L, w, (hipStream_t)stream, nullptr, tape_x,
                             (double *)tape_v);
}

int cd_step(const cd_agents_t *a, const cd_step_params_t *p, const cd_workspace_t *w,
            cd_stream_t stream) {
  return cd_step_tape(a, p, w, nullptr, nullptr, stream);
}

int cd_batch_prepare(const cd_agents_t *a, const cd_step_params_t *p,
                     const cd_scenario_t *scenarios, int32_t n_scenarios, cd_workspace_t *w,
                     uint64_t first_step, cd_stream_t stream) {
  int rc = check_step(a, p, w);
  if (rc) return rc;
  if (!scenarios || n_scenarios < 1 || n_scenarios > w->max_scenarios) return CD_ERR_ARG;
  HostBatch H;
  if ((rc = build_batch(scenarios, n_scenarios, w, p->n_segments, H))) return rc;
  const Layout Lw = make_layout(w);
  char *b = (char *)w->ptr;
  hipStream_t s = (hipStream_t)stream;
  if (hipMemcpyAsync(b + Lw.scn_desc, H.desc.data(), sizeof(ScnDesc) * H.desc.size(),
                     hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(b + Lw.map_f, H.map_f.data(), sizeof(int2) * H.map_f.size(),
                     hipMemcpyHostToDevice, s) != hipSuccess ||
      (!H.map_e.empty() &&
       hipMemcpyAsync(b + Lw.map_e, H.map_e.data(), sizeof(int2) * H.map_e.size(),
                      hipMemcpyHostToDevice, s) != hipSuccess) ||
      hipMemcpyAsync(b + Lw.map_t, H.map_t.data(), sizeof(int2) * H.map_t.size(),
                     hipMemcpyHostToDevice, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return CD_ERR_HIP;
  w->n_scenarios = n_scenarios;
  w->blocks_f = (int32_t)H.map_f.size();
  w->blocks_e = (int32_t)H.map_e.size();
  w->blocks_t = (int32_t)H.map_t.size();
  const Work W = carve(w);
  const Batch B = batch_of(w);
  hipLaunchKernelGGL(k_set_step, dim3(1), dim3(1), 0, s, W.st, (unsigned long long)first_step);
  if (a->n > 0)
    hipLaunchKernelGGL(k_order_identity, dim3(std::min(grid_for(a->n, kBS), 1024)), dim3(kBS), 0,
                       s, W.order, (int)a->n);
  const int slot = (int)(first_step & 1ull);
  hipLaunchKernelGGL(k_scn_prepare, dim3(grid_for((int64_t)n_scenarios * kAccSub, kBS)),
                     dim3(kBS), 0, s, W, B.sst, (int)n_scenarios, slot);
  if ((rc = check_launch())) return rc;
  if (a->real == CD_F64) return launch_reduce(typed<double>(a), W, B, w->blocks_f, true, s, slot);
  return launch_reduce(typed<float>(a), W, B, w->blocks_f, true, s, slot);
}

int cd_batch_step(const cd_agents_t *a, const cd_step_params_t *p, const cd_workspace_t *w,
                  cd_stream_t stream) {
  int rc = check_step(a, p, w);
  if (rc) return rc;
  if (w->n_scenarios < 1 || w->blocks_f < 1) return CD_ERR_ARG;
  const Work W = carve(w);
  const Batch B = batch_of(w);
  if (a->n == 0) return CD_OK;
  const Launch L = batch_launch(w);
  if (a->real == CD_F64)
    return step_kernels<double>(a, p, W, B, L, w, (hipStream_t)stream, nullptr);
  return step_kernels<float>(a, p, W, B, L, w, (hipStream_t)stream, nullptr);
}

int cd_step_timed(const cd_agents_t *a, const cd_step_params_t *p, const cd_workspace_t *w,
                  int64_t steps, double *kernel_ms, cd_stream_t stream) {
  int rc = check_step(a, p, w);
  if (rc) return rc;
  if (!kernel_ms || steps < 0) return CD_ERR_ARG;
  for (int k = 0; k < 4; ++k) kernel_ms[k] = 0.0;
  if (a->n == 0 || steps == 0) return CD_OK;
  hipStream_t s = (hipStream_t)stream;
  const Work W = carve(w);
  const bool batched = w->n_scenarios >= 1 && w->blocks_f >= 1;
  const Batch B = batched ? batch_of(w)
                          : single_batch(w, (p->flags & CD_USE_OBSTACLE) ? p->n_segments : 0,
                                         p->seed, p->dt_min, p->dt_max);
  const Launch L = batched ? batch_launch(w) : single_launch(w);
  hipEvent_t ev[4];
  for (int k = 0; k < 4; ++k)
    if (hipEventCreate(&ev[k]) != hipSuccess) return CD_ERR_HIP;
  for (int64_t it = 0; it < steps && rc == CD_OK; ++it) {
    rc = a->real == CD_F64 ? step_kernels<double>(a, p, W, B, L, w, s, ev)
                           : step_kernels<float>(a, p, W, B, L, w, s, ev);
    if (rc) break;
    if (hipEventSynchronize(ev[3]) != hipSuccess) {
      rc = CD_ERR_HIP;
      break;
    }
    const int from[4] = {0, 1, 2, 0}, to[4] = {1, 2, 3, 3};
    for (int k = 0; k < 4; ++k) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, ev[from[k]], ev[to[k]]) != hipSuccess) rc = CD_ERR_HIP;
      kernel_ms[k] += ms;
    }
  }
  for (int k = 0; k < 4; ++k) (void)hipEventDestroy(ev[k]);
  return rc;
}

int cd_kernel_timed(const cd_agents_t *a, const cd_step_params_t *p, const cd_workspace_t *w,
                    int32_t kernel, int64_t reps, double *avg_ms, cd_stream_t stream) {
  int rc = check_step(a, p, w);
  if (rc) return rc;
  if (!avg_ms || reps < 1 || kernel < 0 || kernel > 1) return CD_ERR_ARG;
  *avg_ms = 0.0;
  if (a->n == 0) return CD_OK;
  hipStream_t s = (hipStream_t)stream;
  const Work W = carve(w);
  const bool batched = w->n_scenarios >= 1 && w->blocks_f >= 1;
  const Batch B = batched ? batch_of(w)
                          : single_batch(w, (p->flags & CD_USE_OBSTACLE) ? p->n_segments : 0,
                                         p->seed, p->dt_min, p->dt_max);
  const Launch L = batched ? batch_launch(w) : single_launch(w);
  const FusedArgs fa = fused_args(p);
  const int reset = (p->flags & CD_USE_RESET) ? 1 : 0;
  auto bin = [&]() {
    if (a->real == CD_F64)
      launch_step_begin<double>(a, p, W, B, fa, L, s);
    else
      launch_step_begin<float>(a, p, W, B, fa, L, s);
  };
  bin();
  if ((rc = check_launch())) return rc;
  hipEvent_t ev[2];
  for (int k = 0; k < 2; ++k)
    if (hipEventCreate(&ev[k]) != hipSuccess) return CD_ERR_HIP;
  (void)hipEventRecord(ev[0], s);
  for (int64_t r = 0; r < reps && rc == CD_OK; ++r) {
    if (kernel == 0)
      bin();
    else
      rc = a->real == CD_F64 ? launch_forces<double>(a, W, B, fa, true, L, w, s, reset)
                             : launch_forces<float>(a, W, B, fa, true, L, w, s, reset);
    if (rc == CD_OK) rc = check_launch();
  }
  (void)hipEventRecord(ev[1], s);
  float ms = 0.f;
  if (rc == CD_OK && (hipEventSynchronize(ev[1]) != hipSuccess ||
                      hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess))
    rc = CD_ERR_HIP;
  for (int k = 0; k < 2; ++k) (void)hipEventDestroy(ev[k]);
  *avg_ms = (double)ms / (double)reps;
  return rc;
}

int cd_graph_begin(cd_stream_t stream) {
  return hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeRelaxed) == hipSuccess
             ? CD_OK
             : CD_ERR_HIP;
}

int cd_graph_end(cd_stream_t stream, void **graph_exec) {
  if (!graph_exec) return CD_ERR_ARG;
  hipGraph_t g = nullptr;
  if (hipStreamEndCapture((hipStream_t)stream, &g) != hipSuccess || !g) return CD_ERR_HIP;
  hipGraphExec_t e = nullptr;
  const hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (err != hipSuccess) return CD_ERR_HIP;
  *graph_exec = (void *)e;
  return CD_OK;
}

int cd_graph_launch(void *graph_exec, cd_stream_t stream) {
  if (!graph_exec) return CD_ERR_ARG;
  return hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream) == hipSuccess
             ? CD_OK
             : CD_ERR_HIP;
}

int cd_graph_destroy(void *graph_exec) {
  if (!graph_exec) return CD_ERR_ARG;
  return hipGraphExecDestroy((hipGraphExec_t)graph_exec) == hipSuccess ? CD_OK : CD_ERR_HIP;
}

int cd_hypot(const double *xy, double *out, int64_t n, cd_stream_t stream) {
  if (n < 0 || (n > 0 && (!xy || !out)) || n > ((int64_t)1 << 40)) return CD_ERR_ARG;
  if (n == 0) return CD_OK;
  hipLaunchKernelGGL(k_hypot, dim3((unsigned)((n + kBS - 1) / kBS)), dim3(kBS), 0,
                     (hipStream_t)stream, (const double2 *)xy, out, n);
  return check_launch();
}

int cd_philox4x32(const uint32_t *in, uint32_t *out, cd_stream_t stream) {
  if (!in || !out) return CD_ERR_ARG;
  hipLaunchKernelGGL(k_philox, dim3(1), dim3(1), 0, (hipStream_t)stream, in, out);
  return check_launch();
}

const char *cd_version(void) { return "cdamd 0.3 gfx950"; }

}  // extern "C"
