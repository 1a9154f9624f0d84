// MI355X (gfx950) kernels of the crowddynamics per-timestep agent update and
// the C ABI declared in include/cdamd.h.
//
// Data layout in HBM: one contiguous array per agent field (vectors
// interleaved [n][2]), plus a caller-owned workspace that holds the device
// state scalars, the block-list arrays and cell-sorted copies of the fields
// the pair kernel gathers from neighbours.
//
// Binning contract (the reference binds cell_lists, an un-vendored
// dependency, at crowddynamics/core/interactions.py:30,192-195):
//   origin o = per-axis minimum of the positions, l = trunc((p - o) / c),
//   shape = max l + 1, flat cell = l_x * ny + l_y, agents ordered stably by
//   flat cell, and every unordered pair of agents in the same or adjacent
//   cells (|dl_x| <= 1 and |dl_y| <= 1) interacts once.
//
// Pair forces are owner-computes: each agent sums its own side of every
// pair, so there are no atomics on forces and results are run-to-run
// deterministic.

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>

#include "../../include/cdamd.h"
#include "cd_device.h"

using namespace cd;

namespace {

// ---------------------------------------------------------------------------
// Workspace
// ---------------------------------------------------------------------------

enum RedField { R_MINX = 0, R_MINY, R_MAXX, R_MAXY, R_VMAX, R_V0MAX, R_NONFINITE, R_COUNT };

struct DevState {
  unsigned long long red[2][8];  // [0] = current values, [1] = accumulator
  unsigned int done[4];          // last-block counters
  double dt, time_tot;
  double origin[2];
  int nx, ny;
  int valid;  // grid of the last binning fits the capacity
  int C;      // cells of the last binning
  int C_clear;  // cells whose counts are non-zero (to clear before next count)
  int oob;      // an agent fell outside the grid (stale bounding box): binning invalid
  unsigned int scan_epoch;   // tags the scan's tile descriptors of this binning
  unsigned int scan_ticket;  // dynamic tile index of the look-back scan
  int pad1;
  long long status;
  unsigned long long step;
  double v_max, v0_max;
};

constexpr int64_t kAlign = 256;
inline int64_t align_up(int64_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

struct Layout {
  int64_t state, tiles, partials, cell_of, rank, pidx, scell, count, offset, spos, svel, sr;
  int64_t sls, srs, sphi, srt, srs_r, srts;
  int64_t total;
};

inline int64_t partials_bytes(int64_t n) {
  return 8 * 8 * (std::max<int64_t>(512, (n + 255) / 256) + 1);
}

Layout make_layout(int64_t n, int64_t cap, int32_t model) {
  Layout L;
  int64_t o = 0;
  auto take = [&](int64_t bytes) {
    int64_t r = o;
    o = align_up(o + bytes);
    return r;
  };
  // state and partials first: their offsets depend on n only
  L.state = take(4096);
  L.partials = take(partials_bytes(n));
  L.tiles = take(8 * ((cap + kScanTile - 1) / kScanTile + 1));
  L.cell_of = take(4 * n);
  L.rank = take(4 * n);
  L.pidx = take(4 * n);
  L.scell = take(4 * n);
  L.count = take(4 * cap);
  L.offset = take(4 * (cap + 1));
  L.spos = take(16 * n);
  L.svel = take(16 * n);
  L.sr = take(8 * n);
  if (model == CD_THREE_CIRCLE) {
    L.sls = take(16 * n);
    L.srs = take(16 * n);
    L.sphi = take(8 * n);
    L.srt = take(8 * n);
    L.srs_r = take(8 * n);
    L.srts = take(8 * n);
  } else {
    L.sls = L.srs = L.sphi = L.srt = L.srs_r = L.srts = 0;
  }
  L.total = o;
  return L;
}

template <typename T>
struct Agents {
  int n;
  double2 *pos;
  V2<T> *vel, *e0, *force, *force_prev;
  double2 *ls, *rs;
  const double *radius, *r_t, *r_s, *r_ts, *mass, *inertia, *v0, *w0;
  const double *tau_adj, *k_soc, *tau_0, *mu, *kappa, *damping, *sig_f;
  double *phi, *omega, *phi0, *torque, *torque_prev;
  const double *tau_rot, *sig_t;
  const int64_t *target;
};

template <typename T>
Agents<T> typed(const cd_agents_t *a) {
  Agents<T> A;
  A.n = (int)a->n;
  A.pos = (double2 *)a->position;
  A.vel = (V2<T> *)a->velocity;
  A.e0 = (V2<T> *)a->target_direction;
  A.force = (V2<T> *)a->force;
  A.force_prev = (V2<T> *)a->force_prev;
  A.ls = (double2 *)a->position_ls;
  A.rs = (double2 *)a->position_rs;
  A.radius = a->radius;
  A.r_t = a->r_t;
  A.r_s = a->r_s;
  A.r_ts = a->r_ts;
  A.mass = a->mass;
  A.inertia = a->inertia_rot;
  A.v0 = a->target_velocity;
  A.w0 = a->target_angular_velocity;
  A.tau_adj = a->tau_adj;
  A.k_soc = a->k_soc;
  A.tau_0 = a->tau_0;
  A.mu = a->mu;
  A.kappa = a->kappa;
  A.damping = a->damping;
  A.sig_f = a->std_rand_force;
  A.phi = a->orientation;
  A.omega = a->angular_velocity;
  A.phi0 = a->target_orientation;
  A.torque = a->torque;
  A.torque_prev = a->torque_prev;
  A.tau_rot = a->tau_rot;
  A.sig_t = a->std_rand_torque;
  A.target = a->target;
  return A;
}

struct Work {
  DevState *st;
  unsigned long long *tiles;
  unsigned long long *partials;
  int ntiles;
  int *cell_of, *rank, *pidx, *scell, *count, *offset;
  double2 *spos;
  void *svel;
  double *sr;
  double2 *sls, *srs;
  double *sphi, *srt, *srs_r, *srts;
  int cap;
};

Work carve(void *ws, const Layout &L, int64_t cap) {
  char *b = (char *)ws;
  Work W;
  W.st = (DevState *)(b + L.state);
  W.tiles = (unsigned long long *)(b + L.tiles);
  W.partials = (unsigned long long *)(b + L.partials);
  W.ntiles = (int)((cap + kScanTile - 1) / kScanTile);
  W.cell_of = (int *)(b + L.cell_of);
  W.rank = (int *)(b + L.rank);
  W.pidx = (int *)(b + L.pidx);
  W.scell = (int *)(b + L.scell);
  W.count = (int *)(b + L.count);
  W.offset = (int *)(b + L.offset);
  W.spos = (double2 *)(b + L.spos);
  W.svel = (void *)(b + L.svel);
  W.sr = (double *)(b + L.sr);
  W.sls = L.sls ? (double2 *)(b + L.sls) : nullptr;
  W.srs = L.srs ? (double2 *)(b + L.srs) : nullptr;
  W.sphi = L.sphi ? (double *)(b + L.sphi) : nullptr;
  W.srt = L.srt ? (double *)(b + L.srt) : nullptr;
  W.srs_r = L.srs_r ? (double *)(b + L.srs_r) : nullptr;
  W.srts = L.srts ? (double *)(b + L.srts) : nullptr;
  W.cap = (int)cap;
  return W;
}

constexpr unsigned long long kOrdPosInf = 0xFFF0000000000000ull;  // ord_bits(+inf)
constexpr unsigned long long kOrdNegInf = 0x000FFFFFFFFFFFFFull;  // ord_bits(-inf)

__device__ __forceinline__ void reset_acc(DevState *st) {
  st->red[1][R_MINX] = kOrdPosInf;
  st->red[1][R_MINY] = kOrdPosInf;
  st->red[1][R_MAXX] = kOrdNegInf;
  st->red[1][R_MAXY] = kOrdNegInf;
  st->red[1][R_VMAX] = 0ull;
  st->red[1][R_V0MAX] = kOrdNegInf;
  st->red[1][R_NONFINITE] = 0ull;
  st->red[1][7] = 0ull;
}

__global__ void __launch_bounds__(256) k_ws_init(DevState *st, int *count, int cap,
                                                 unsigned long long *tiles, int ntiles) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  for (int k = t; k < ntiles; k += gridDim.x * blockDim.x) tiles[k] = 0ull;
  if (t == 0) {
    memset(st, 0, sizeof(DevState));
    for (int k = 0; k < 8; ++k) st->red[0][k] = 0ull;
    reset_acc(st);
  }
  for (int c = t; c < cap; c += gridDim.x * blockDim.x) count[c] = 0;
}

// Block-wide reduction helpers (wave64 shuffles, then LDS across waves).
template <typename F>
__device__ __forceinline__ unsigned long long wave_reduce(unsigned long long v, F op) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    unsigned long long w = __shfl_xor(v, o, 64);
    v = op(v, w);
  }
  return v;
}

struct MinOp {
  __device__ unsigned long long operator()(unsigned long long a, unsigned long long b) const {
    return a < b ? a : b;
  }
};
struct MaxOp {
  __device__ unsigned long long operator()(unsigned long long a, unsigned long long b) const {
    return a > b ? a : b;
  }
};

__device__ __forceinline__ void combine(unsigned long long *v, const unsigned long long *u) {
  v[R_MINX] = min(v[R_MINX], u[R_MINX]);
  v[R_MINY] = min(v[R_MINY], u[R_MINY]);
  v[R_MAXX] = max(v[R_MAXX], u[R_MAXX]);
  v[R_MAXY] = max(v[R_MAXY], u[R_MAXY]);
  v[R_VMAX] = max(v[R_VMAX], u[R_VMAX]);
  v[R_V0MAX] = max(v[R_V0MAX], u[R_V0MAX]);
  v[R_NONFINITE] = max(v[R_NONFINITE], u[R_NONFINITE]);
}

// Block reduction of the per-thread values into thread 0's v.
__device__ void block_reduce(unsigned long long v[R_COUNT]) {
  __shared__ unsigned long long sh[R_COUNT][16];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  v[R_MINX] = wave_reduce(v[R_MINX], MinOp());
  v[R_MINY] = wave_reduce(v[R_MINY], MinOp());
  v[R_MAXX] = wave_reduce(v[R_MAXX], MaxOp());
  v[R_MAXY] = wave_reduce(v[R_MAXY], MaxOp());
  v[R_VMAX] = wave_reduce(v[R_VMAX], MaxOp());
  v[R_V0MAX] = wave_reduce(v[R_V0MAX], MaxOp());
  v[R_NONFINITE] = wave_reduce(v[R_NONFINITE], MaxOp());
  __syncthreads();
  if (lane == 0)
    for (int k = 0; k < R_COUNT; ++k) sh[k][wid] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < nw; ++w) {
      unsigned long long u[R_COUNT];
      for (int k = 0; k < R_COUNT; ++k) u[k] = sh[k][w];
      combine(v, u);
    }
  }
}

// Reduces this block's values, stores them as the block's partial, and lets
// the last block to finish reduce all partials into the current slot
// (st->red[0]).  One atomic per block (the completion counter); the
// partials are written and read with agent-scope atomics so every XCD sees
// them.  Returns true in the last block.
__device__ bool block_reduce_commit(DevState *st, unsigned long long v[R_COUNT], int counter,
                                    unsigned long long *partials) {
  __shared__ bool last;
  block_reduce(v);
  if (threadIdx.x == 0) {
    for (int k = 0; k < R_COUNT; ++k)
      __hip_atomic_store(&partials[blockIdx.x * 8 + k], v[k], __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    unsigned int prev = atomicAdd(&st->done[counter], 1u);
    last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
  unsigned long long w[R_COUNT] = {kOrdPosInf, kOrdPosInf, kOrdNegInf, kOrdNegInf, 0ull,
                                   kOrdNegInf, 0ull};
  for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) {
    unsigned long long u[R_COUNT];
    for (int k = 0; k < R_COUNT; ++k)
      u[k] = __hip_atomic_load(&partials[b * 8 + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    combine(w, u);
  }
  block_reduce(w);
  if (threadIdx.x == 0) {
    for (int k = 0; k < R_COUNT; ++k) st->red[0][k] = w[k];
    st->done[counter] = 0u;
    st->oob = 0;
    if (w[R_NONFINITE]) st->status |= CD_STATUS_NONFINITE;
  }
  return true;
}

// Reduction of the binning bounding box, max |v| (hypot, strict >, NaN
// ignored as in core/integrator.py:81-85) and max target_velocity; also
// clears the counts of the previous binning.
template <typename T>
__global__ void __launch_bounds__(256) k_reduce(Agents<T> A, DevState *st, int *count,
                                                unsigned long long *partials) {
  const int n = A.n;
  if (count) {
    const int clear = st->C_clear;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < clear; c += gridDim.x * blockDim.x)
      count[c] = 0;
  }
  unsigned long long v[R_COUNT] = {kOrdPosInf, kOrdPosInf, kOrdNegInf, kOrdNegInf, 0ull,
                                   kOrdNegInf, 0ull};
  double vmax = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double2 p = A.pos[i];
    if (!(isfinite(p.x) && isfinite(p.y))) v[R_NONFINITE] = 1ull;
    unsigned long long bx = ord_bits(p.x), by = ord_bits(p.y);
    v[R_MINX] = min(v[R_MINX], bx);
    v[R_MINY] = min(v[R_MINY], by);
    v[R_MAXX] = max(v[R_MAXX], bx);
    v[R_MAXY] = max(v[R_MAXY], by);
    V2<T> w = A.vel[i];
    double l = dhypot((double)w.x, (double)w.y);
    if (l > vmax) vmax = l;
    v[R_V0MAX] = max(v[R_V0MAX], ord_bits(A.v0[i]));
  }
  v[R_VMAX] = (unsigned long long)__double_as_longlong(vmax);
  if (block_reduce_commit(st, v, 0, partials) && threadIdx.x == 0 && count) st->C_clear = 0;
}

// Grid of the binning from the current bounding box (contract in the file
// header); all threads compute the same values.
struct Grid {
  double ox, oy;
  int nx, ny;
  bool ok;
};

__device__ __forceinline__ Grid grid_from_state(const DevState *st, double cell, int cap) {
  Grid g;
  double mnx = from_ord_bits(st->red[0][R_MINX]), mny = from_ord_bits(st->red[0][R_MINY]);
  double mxx = from_ord_bits(st->red[0][R_MAXX]), mxy = from_ord_bits(st->red[0][R_MAXY]);
  g.ox = mnx;
  g.oy = mny;
  double qx = (mxx - mnx) / cell, qy = (mxy - mny) / cell;
  g.ok = st->red[0][R_NONFINITE] == 0ull && qx >= 0.0 && qy >= 0.0 && qx < (double)cap &&
         qy < (double)cap;
  if (g.ok) {
    long long nx = (long long)qx + 1, ny = (long long)qy + 1;
    g.ok = nx * ny <= (long long)cap;
    g.nx = (int)nx;
    g.ny = (int)ny;
  } else {
    g.nx = g.ny = 0;
  }
  return g;
}

// K1a: cell of each agent and its (arbitrary) rank within the cell.
__global__ void __launch_bounds__(256) k_bin_count(const double2 *pos, int n, double cell, DevState *st, int *count,
                            int *cell_of, int *rank, int cap) {
  Grid g = grid_from_state(st, cell, cap);
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    st->origin[0] = g.ox;
    st->origin[1] = g.oy;
    st->nx = g.nx;
    st->ny = g.ny;
    st->valid = g.ok ? 1 : 0;
    st->C = g.ok ? g.nx * g.ny : 0;
    st->C_clear = g.ok ? g.nx * g.ny : 0;
    st->scan_epoch = (st->scan_epoch + 1u) & 0x3fffffffu;
    if (st->scan_epoch == 0u) st->scan_epoch = 1u;
    st->scan_ticket = 0u;
    if (!g.ok) st->status |= CD_STATUS_GRID_OVERFLOW;
  }
  if (!g.ok || i >= n) return;
  double2 p = pos[i];
  double qx = (p.x - g.ox) / cell, qy = (p.y - g.oy) / cell;
  // positions changed after the bounding box was reduced: refuse to bin
  if (!(qx >= 0.0 && qy >= 0.0 && qx < (double)g.nx && qy < (double)g.ny)) {
    atomicOr(&st->oob, 1);
    return;
  }
  int lx = (int)qx;
  int ly = (int)qy;
  int c = lx * g.ny + ly;
  cell_of[i] = c;
  rank[i] = atomicAdd(&count[c], 1);
}

// K1b: exclusive scan of the counts, single pass with decoupled look-back.
// Each workgroup takes the next tile of kScanTile cells (dynamic ticket, so
// every tile's predecessors are already resident), publishes its aggregate,
// then walks back over predecessors' descriptors until an inclusive prefix.
// A descriptor packs (epoch:30 | flag:2 | value:32); the epoch of this
// binning makes stale descriptors of earlier steps invisible.  The wait is
// bounded: past kSpinLimit polls the binning is marked invalid instead.
constexpr long long kSpinLimit = 1ll << 24;

__device__ __forceinline__ unsigned long long scan_pack(unsigned epoch, unsigned flag, int v) {
  return ((unsigned long long)epoch << 34) | ((unsigned long long)flag << 32) | (unsigned)v;
}

__global__ void __launch_bounds__(kScanThreads) k_bin_scan(DevState *st, const int *count,
                                                           int *offset,
                                                           unsigned long long *tiles) {
  __shared__ int s_tile, s_prefix;
  __shared__ int s_wave[kScanThreads / 64];
  if (st->oob) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->valid = 0;
      st->status |= CD_STATUS_GRID_OVERFLOW;
    }
    return;
  }
  if (!st->valid) return;
  const int C = st->C;
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
    int idx = base + k;
    v[k] = idx < C ? count[idx] : 0;
    sum += v[k];
  }
  // inclusive scan of the thread sums within each wave, then across waves
  int inc = sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int u = __shfl_up(inc, o, 64);
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
  if (wid == 0) {
    // publish this tile's aggregate, then look back over up to 64 predecessor
    // descriptors at once (one per lane) until an inclusive prefix is found
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
        unsigned long long d =
            __hip_atomic_load(&tiles[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        flag = ((unsigned)(d >> 34) == epoch) ? ((unsigned)(d >> 32) & 3u) : 0u;
        val = (int)(unsigned)d;
      }
      const unsigned long long incl = __ballot(flag == 2u);
      const unsigned long long bad = __ballot(flag == 0u);
      // lanes up to (and including) the nearest inclusive descriptor count
      const int stop = incl ? (__ffsll((long long)incl) - 1) : 63;
      const unsigned long long need =
          (stop >= 63) ? ~0ull : ((2ull << stop) - 1ull);
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
        st->valid = 0;
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
    int idx = base + k;
    if (idx < C) offset[idx] = run;
    run += v[k];
  }
  if (tile == ntiles - 1 && t == kScanThreads - 1) offset[C] = s_prefix + agg;
}

// K1c: scatter agents to their cell's segment.
__global__ void __launch_bounds__(256) k_bin_scatter(int n, const DevState *st, const int *cell_of, const int *rank,
                              const int *offset, int *pidx) {
  if (!st->valid) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  pidx[offset[cell_of[i]] + rank[i]] = i;
}

// K1d: order each cell's segment by agent index (stable counting sort), then
// gather the cell-sorted copies the pair kernel reads.
template <typename T>
__global__ void __launch_bounds__(256) k_bin_sort_gather(Agents<T> A, const DevState *st, const int *offset, int *pidx,
                                  int *scell, double2 *spos, V2<T> *svel, double *sr,
                                  double2 *sls, double2 *srs, double *sphi, double *srt,
                                  double *srs_r, double *srts) {
  if (!st->valid) return;
  const int C = st->C;
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  const int s0 = offset[c], s1 = offset[c + 1];
  if (s1 - s0 > 1) {
    for (int s = s0 + 1; s < s1; ++s) {
      int key = pidx[s];
      int t = s - 1;
      while (t >= s0 && pidx[t] > key) {
        pidx[t + 1] = pidx[t];
        --t;
      }
      pidx[t + 1] = key;
    }
  }
  for (int s = s0; s < s1; ++s) {
    int i = pidx[s];
    scell[s] = c;
    spos[s] = A.pos[i];
    svel[s] = A.vel[i];
    sr[s] = A.radius[i];
    if (sls) {
      sls[s] = A.ls[i];
      srs[s] = A.rs[i];
      sphi[s] = A.phi[i];
      srt[s] = A.r_t[i];
      srs_r[s] = A.r_s[i];
      srts[s] = A.r_ts[i];
    }
  }
}

// ---------------------------------------------------------------------------
// Per-agent operators
// ---------------------------------------------------------------------------

// Fluctuation noise of agent i (core/motion/fluctuation.py:14-62).
__device__ __forceinline__ void fluct_noise(uint64_t seed, uint64_t step, int i, double mass,
                                            double sig_f, double inertia, double sig_t, double &fx,
                                            double &fy, double &m) {
  double ua, ub;
  draw2(seed, i, step, 0, 0, ua, ub);
  double phi = 2.0 * kPi * ua;
  double mag = sig_f * fabs(truncnorm3(seed, i, step, 1));
  double mm = mass * mag;
  fx = mm * cos(phi);
  fy = mm * sin(phi);
  m = inertia * (sig_t * truncnorm3(seed, i, step, 2));
}

template <typename T>
__global__ void __launch_bounds__(256) k_fluct(Agents<T> A, int model, uint64_t seed, uint64_t step) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  double fx, fy, m;
  fluct_noise(seed, step, i, A.mass[i], A.sig_f[i], model ? A.inertia[i] : 0.0,
              model ? A.sig_t[i] : 0.0, fx, fy, m);
  V2<T> f = A.force[i];
  f.x += (T)fx;
  f.y += (T)fy;
  A.force[i] = f;
  if (model) A.torque[i] += m;
}

// Navigation lookup of agent i; returns false when e0 is kept.
__device__ __forceinline__ bool nav_lookup(double px, double py, int64_t t, const double *maps,
                                           int64_t T_, int64_t rows, int64_t cols, double minx,
                                           double miny, double step, int mode, double &ex,
                                           double &ey) {
  if (t < 0 || t >= T_) return false;
  double qx = (px - minx) / step;
  double qy = (py - miny) / step;
  if (mode == 0) {
    // indicer truncation + is_inside (quickest_path.py:41-44, navigation.py:59-78)
    if (!(qx > -1.0 && qx < (double)cols && qy > -1.0 && qy < (double)rows)) return false;
    int64_t jj = (int64_t)qx, ii = (int64_t)qy;
    const double2 m = ((const double2 *)maps)[(t * rows + ii) * cols + jj];
    ex = m.x;
    ey = m.y;
    return true;
  }
  if (!(qx >= 0.0 && qx <= (double)(cols - 1) && qy >= 0.0 && qy <= (double)(rows - 1)))
    return false;
  int64_t c0 = (int64_t)qx, r0 = (int64_t)qy;
  if (c0 > cols - 2) c0 = cols - 2;
  if (r0 > rows - 2) r0 = rows - 2;
  if (c0 < 0 || r0 < 0) return false;
  double fx = qx - (double)c0, fy = qy - (double)r0;
  const double *P = maps + t * rows * cols;
  double p00 = P[r0 * cols + c0], p01 = P[r0 * cols + c0 + 1];
  double p10 = P[(r0 + 1) * cols + c0], p11 = P[(r0 + 1) * cols + c0 + 1];
  double gx = ((p01 - p00) * (1.0 - fy) + (p11 - p10) * fy) / step;
  double gy = ((p10 - p00) * (1.0 - fx) + (p11 - p01) * fx) / step;
  double g = hypot(gx, gy);
  if (!(g > 0.0)) return false;
  ex = -gx / g;
  ey = -gy / g;
  return true;
}

template <typename T>
__global__ void __launch_bounds__(256) k_nav(Agents<T> A, const double *maps, int64_t T_, int64_t rows, int64_t cols,
                      double minx, double miny, double step, int mode) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  double2 p = A.pos[i];
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
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  V2<T> e = A.e0[i];
  A.phi0[i] = atan2((double)e.y, (double)e.x);  // orientation.py:21
}

// force_adjust (core/motion/adjusting.py:16-51)
template <typename T>
__device__ __forceinline__ void adjust_force(const Agents<T> &A, int i, V2<T> v, V2<T> e,
                                             T &fx, T &fy) {
  T s = (T)(A.mass[i] / A.tau_adj[i]);
  T v0 = (T)A.v0[i];
  fx = s * (v0 * e.x - v.x);
  fy = s * (v0 * e.y - v.y);
}

// torque_adjust (core/motion/adjusting.py:54-95)
template <typename T>
__device__ __forceinline__ double adjust_torque(const Agents<T> &A, int i) {
  double s = A.inertia[i] / A.tau_rot[i];
  return s * ((wrap_to_pi(A.phi0[i] - A.phi[i]) / kPi) * A.w0[i] - A.omega[i]);
}

template <typename T>
__global__ void __launch_bounds__(256) k_adjust(Agents<T> A, int do_force, int do_torque) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
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
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  V2<T> z;
  z.x = T(0);
  z.y = T(0);
  A.force[i] = z;
  if (model) A.torque[i] = 0.0;
}

// Wall contacts of agent i in segment order (core/interactions.py:107-141,
// 169-186).  Segments are staged through LDS by the caller.
template <typename T>
__device__ __forceinline__ void walls_agent(const Agents<T> &A, int model, int i, double2 p,
                                            V2<T> v, const double4 *seg, int W, T &fx, T &fy,
                                            double &tq, int w_begin = 0, int w_stride = 1) {
  const double r = A.radius[i];
  double2 ls, rs;
  double rt = 0.0, rsr = 0.0;
  if (model) {
    ls = A.ls[i];
    rs = A.rs[i];
    rt = A.r_t[i];
    rsr = A.r_s[i];
  }
  const T mu = (T)A.mu[i], kappa = (T)A.kappa[i], damping = (T)A.damping[i];
  for (int w = w_begin; w < W; w += w_stride) {
    double4 s = seg[w];
    double h, nx, ny, rmx = 0.0, rmy = 0.0;
    if (!model) {
      distance_circle_line(p.x, p.y, r, s.x, s.y, s.z, s.w, h, nx, ny);
    } else {
      // distance_three_circle_line, core/distance.py:149-180
      const double px[3] = {p.x, ls.x, rs.x}, py[3] = {p.y, ls.y, rs.y};
      const double rr[3] = {rt, rsr, rsr};
      double h_min = __builtin_nan("");
      double bx = 0.0, by = 0.0, sx = 0.0, sy = 0.0, sr_ = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double hk, ex, ey;
        distance_circle_line(px[k], py[k], rr[k], s.x, s.y, s.z, s.w, hk, ex, ey);
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
      T tnx = (T)nx, tny = (T)ny;
      force_contact((T)h, tnx, tny, v.x, v.y, tny, -tnx, mu, kappa, damping, cx, cy);
      fx += cx;
      fy += cy;
      if (model) tq += rmx * (double)cy - rmy * (double)cx;
    }
  }
}

constexpr int kSegTile = 512;

template <typename T>
__global__ void __launch_bounds__(256) k_walls(Agents<T> A, int model, const double4 *segs, int W) {
  __shared__ double4 sseg[kSegTile];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool live = i < A.n;
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
    int cnt = min(kSegTile, W - w0);
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
// Pair forces (owner-computes over the 3x3 cell stencil)
// ---------------------------------------------------------------------------
//
// Four lanes serve one agent (the agent of sorted slot s = thread / 4).
// Lanes 0..2 sum the agent's side of every pair in stencil row dx = -1, 0, +1
// (the three cells of a row are contiguous in the cell-sorted order, so a
// row is one slot range); lane 3 handles the agent's own terms
// (fluctuation, navigation, orientation, adjusting) and the wall contacts
// are split over the four lanes.  The partial sums meet on lane 3 through
// wavefront shuffles in a fixed order, so results are deterministic.

struct Slots {
  const int *offset, *pidx, *scell;
  const double2 *spos;
  const void *svel;
  const double *sr;
  const double2 *sls, *srs;
  const double *sphi, *srt, *srs_r, *srts;
};

Slots slots_of(const Work &W) {
  Slots S;
  S.offset = W.offset;
  S.pidx = W.pidx;
  S.scell = W.scell;
  S.spos = W.spos;
  S.svel = W.svel;
  S.sr = W.sr;
  S.sls = W.sls;
  S.srs = W.srs;
  S.sphi = W.sphi;
  S.srt = W.srt;
  S.srs_r = W.srs_r;
  S.srts = W.srts;
  return S;
}

template <typename T>
__device__ __forceinline__ Owner<T> owner_of(const Agents<T> &A, int i) {
  Owner<T> o;
  o.neg_mk = (T)((-A.mass[i]) * A.k_soc[i]);
  o.tau_0 = (T)A.tau_0[i];
  o.mu = (T)A.mu[i];
  o.kappa = (T)A.kappa[i];
  o.damping = (T)A.damping[i];
  return o;
}

// Slot range of stencil row ox (cells ylo..yhi of column ox).
__device__ __forceinline__ void row_range(const int *offset, int ox, int ny, int ly, int &k0,
                                          int &k1) {
  const int ylo = max(ly - 1, 0), yhi = min(ly + 1, ny - 1);
  k0 = offset[ox * ny + ylo];
  k1 = offset[ox * ny + yhi + 1];
}

// Circular pairs of one stencil row, neighbour loads pipelined one ahead.
template <typename T>
__device__ __forceinline__ void row_circular(const Slots &S, int k0, int k1, int s, double2 xi,
                                             V2<T> vi, double ri, const Owner<T> &o, T &fx,
                                             T &fy) {
  const V2<T> *svel = (const V2<T> *)S.svel;
  if (k0 >= k1) return;
  double2 xn = S.spos[k0];
  V2<T> vn = svel[k0];
  double rn = S.sr[k0];
  for (int k = k0; k < k1; ++k) {
    const double2 xj = xn;
    const V2<T> vj = vn;
    const double rj = rn;
    if (k + 1 < k1) {
      xn = S.spos[k + 1];
      vn = svel[k + 1];
      rn = S.sr[k + 1];
    }
    if (k == s) continue;
    T px, py;
    if (pair_circular<T>((T)(xi.x - xj.x), (T)(xi.y - xj.y), vi.x - vj.x, vi.y - vj.y,
                         (T)(ri + rj), o, px, py)) {
      fx += px;
      fy += py;
    }
  }
}

__device__ __forceinline__ TCView tc_view(const Slots &S, int k) {
  TCView t;
  double2 p = S.spos[k], l = S.sls[k], r = S.srs[k];
  t.px[0] = p.x;
  t.py[0] = p.y;
  t.px[1] = l.x;
  t.py[1] = l.y;
  t.px[2] = r.x;
  t.py[2] = r.y;
  V2<double> v = ((const V2<double> *)S.svel)[k];
  t.vx = v.x;
  t.vy = v.y;
  t.r_t = S.srt[k];
  t.r_s = S.srs_r[k];
  t.r_ts = S.srts[k];
  t.phi = S.sphi[k];
  return t;
}

// Three-circle pairs of one stencil row: each pair is evaluated in the
// reference's (min index, max index) argument order, keeping the owner's side.
__device__ __forceinline__ void row_three_circle(const Slots &S, int k0, int k1, int s, int i,
                                                 const TCView &me, const Owner<double> &o,
                                                 double &fx, double &fy, double &tq) {
  for (int k = k0; k < k1; ++k) {
    if (k == s) continue;
    const int j = S.pidx[k];
    const TCView other = tc_view(S, k);
    double px, py, m;
    bool hit = (i < j) ? pair_three_circle(me, other, false, o, px, py, m)
                       : pair_three_circle(other, me, true, o, px, py, m);
    if (hit) {
      fx += px;
      fy += py;
      tq += m;
    }
  }
}

struct FusedArgs {
  int flags, nav_mode;
  uint64_t seed;
  const double *maps;
  int64_t nav_T, nav_rows, nav_cols;
  double nav_minx, nav_miny, nav_step;
  const double4 *segs;
  int W;
};

// FUSED = false: the AgentAgentInteractions node, force (torque) += pairs.
// FUSED = true: all per-agent forces of one step in the post-order of the
// standard logic tree (fluctuation, navigation, orientation, adjusting,
// pairs, walls), written to force (torque).
template <typename T, int MODEL, bool FUSED>
__global__ void __launch_bounds__(256) k_forces(Agents<T> A, DevState *st, FusedArgs fa,
                                                Slots S) {
  __shared__ double4 sseg[kSegTile];
  const int gt = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = gt >> 2;
  const int q = threadIdx.x & 3;
  const int lane0 = (threadIdx.x & 63) & ~3;
  const bool pairs_on = FUSED ? ((fa.flags & CD_USE_AGENT_AGENT) != 0) : true;
  const bool sorted = pairs_on && st->valid;
  if (!FUSED && !sorted) return;  // node mode: binning invalid, nothing to add
  const bool have = s < A.n;
  int i = -1, c = 0;
  if (have) {
    if (sorted) {
      i = S.pidx[s];
      c = S.scell[s];
    } else {
      i = s;
    }
  }
  T px = T(0), py = T(0);  // this lane's pair partial sum
  T wx = T(0), wy = T(0);  // this lane's wall partial sum
  double ptq = 0.0, wtq = 0.0;
  double2 pos;
  V2<T> vel;
  if (have) {
    pos = sorted ? S.spos[s] : A.pos[i];
    vel = sorted ? ((const V2<T> *)S.svel)[s] : A.vel[i];
  }
  if (have && sorted && q < 3) {
    const int ny = st->ny, nx = st->nx;
    const int lx = c / ny, ly = c - lx * ny;
    const int ox = lx + q - 1;
    if (ox >= 0 && ox < nx) {
      int k0, k1;
      row_range(S.offset, ox, ny, ly, k0, k1);
      if constexpr (MODEL == 0) {
        row_circular<T>(S, k0, k1, s, pos, vel, S.sr[s], owner_of(A, i), px, py);
      } else {
        if constexpr (sizeof(T) == 8) {
          double fx = 0.0, fy = 0.0;
          row_three_circle(S, k0, k1, s, i, tc_view(S, s), owner_of(A, i), fx, fy, ptq);
          px = fx;
          py = fy;
        }
      }
    }
  }
  if (FUSED && (fa.flags & CD_USE_OBSTACLE)) {
    for (int w0 = 0; w0 < fa.W; w0 += kSegTile) {
      const int cnt = min(kSegTile, fa.W - w0);
      __syncthreads();
      for (int k = threadIdx.x; k < cnt; k += blockDim.x) sseg[k] = fa.segs[w0 + k];
      __syncthreads();
      if (have) walls_agent(A, MODEL, i, pos, vel, sseg, cnt, wx, wy, wtq, q, 4);
    }
  }
  // partial sums of lanes 0..3 to lane 3 (all lanes take part in the shuffles)
  const T p0x = __shfl(px, lane0 + 0, 64), p0y = __shfl(py, lane0 + 0, 64);
  const T p1x = __shfl(px, lane0 + 1, 64), p1y = __shfl(py, lane0 + 1, 64);
  const T p2x = __shfl(px, lane0 + 2, 64), p2y = __shfl(py, lane0 + 2, 64);
  const double t0 = __shfl(ptq, lane0 + 0, 64), t1 = __shfl(ptq, lane0 + 1, 64),
               t2 = __shfl(ptq, lane0 + 2, 64);
  T w0x = T(0), w0y = T(0), w1x = T(0), w1y = T(0), w2x = T(0), w2y = T(0);
  double u0 = 0.0, u1 = 0.0, u2 = 0.0;
  if (FUSED) {
    w0x = __shfl(wx, lane0 + 0, 64);
    w0y = __shfl(wy, lane0 + 0, 64);
    w1x = __shfl(wx, lane0 + 1, 64);
    w1y = __shfl(wy, lane0 + 1, 64);
    w2x = __shfl(wx, lane0 + 2, 64);
    w2y = __shfl(wy, lane0 + 2, 64);
    u0 = __shfl(wtq, lane0 + 0, 64);
    u1 = __shfl(wtq, lane0 + 1, 64);
    u2 = __shfl(wtq, lane0 + 2, 64);
  }
  if (!have || q != 3) return;
  V2<T> f = A.force[i];
  double tq = MODEL ? A.torque[i] : 0.0;
  if (FUSED) {
    if (fa.flags & CD_USE_FLUCT) {
      double nfx, nfy, nm = 0.0;
      double ua, ub;
      draw2(fa.seed, i, st->step, 0, 0, ua, ub);
      const double phi = 2.0 * kPi * ua;
      const double mm = A.mass[i] * (A.sig_f[i] * fabs(truncnorm3(fa.seed, i, st->step, 1)));
      nfx = mm * cos(phi);
      nfy = mm * sin(phi);
      if (MODEL) nm = A.inertia[i] * (A.sig_t[i] * truncnorm3(fa.seed, i, st->step, 2));
      f.x += (T)nfx;
      f.y += (T)nfy;
      if (MODEL) tq += nm;
    }
    V2<T> e = A.e0[i];
    if (fa.flags & CD_USE_NAV) {
      double ex, ey;
      if (nav_lookup(pos.x, pos.y, A.target[i], fa.maps, fa.nav_T, fa.nav_rows, fa.nav_cols,
                     fa.nav_minx, fa.nav_miny, fa.nav_step, fa.nav_mode, ex, ey)) {
        e.x = (T)ex;
        e.y = (T)ey;
        A.e0[i] = e;
      }
    }
    if (fa.flags & CD_USE_ADJUST) {
      if (MODEL) A.phi0[i] = atan2((double)e.y, (double)e.x);
      T ax, ay;
      adjust_force(A, i, vel, e, ax, ay);
      f.x += ax;
      f.y += ay;
      if (MODEL) tq += adjust_torque(A, i);
    }
  }
  f.x += (p0x + p1x) + p2x;
  f.y += (p0y + p1y) + p2y;
  if (MODEL) tq += (t0 + t1) + t2;
  if (FUSED) {
    f.x += (w0x + w1x) + (w2x + wx);
    f.y += (w0y + w1y) + (w2y + wy);
    if (MODEL) tq += (u0 + u1) + (u2 + wtq);
  }
  A.force[i] = f;
  if (MODEL) A.torque[i] = tq;
}

// ---------------------------------------------------------------------------
// Integrator
// ---------------------------------------------------------------------------

// adaptive_timestep (core/integrator.py:81-97) from the reduced values.
__device__ __forceinline__ double adaptive_dt(const DevState *st, double dt_min, double dt_max) {
  double v_max = __longlong_as_double((long long)st->red[0][R_VMAX]);
  if (v_max == 0.0) return dt_max;
  double v0max = from_ord_bits(st->red[0][R_V0MAX]);
  double dx_max = 1.1 * v0max * dt_max;
  double dt = dx_max / v_max;
  if (dt > dt_max) return dt_max;
  if (dt < dt_min) return dt_min;
  return dt;
}

// translational_verlet (core/integrator.py:164-176), rotational_verlet
// (:179-193) and shoulders (simulation/agents.py:473-486).  With `fused`,
// also Reset (simulation/logic.py:59-64), the next step's reductions and the
// clearing of this step's cell counts.
template <typename T>
__global__ void __launch_bounds__(256) k_integrate(Agents<T> A, int model, DevState *st, double dt_min, double dt_max,
                            int fused, int do_reset, int *count,
                            unsigned long long *partials) {
  const double dt = adaptive_dt(st, dt_min, dt_max);
  const int n = A.n;
  unsigned long long v[R_COUNT] = {kOrdPosInf, kOrdPosInf, kOrdNegInf, kOrdNegInf, 0ull,
                                   kOrdNegInf, 0ull};
  double vmax = 0.0;
  if (fused) {
    const int clear = st->C_clear;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < clear; c += gridDim.x * blockDim.x)
      count[c] = 0;
  }
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const double m = A.mass[i];
    V2<T> f = A.force[i], fp = A.force_prev[i], vel = A.vel[i];
    double2 x = A.pos[i];
    const T tm = (T)m;
    T aox = fp.x / tm, aoy = fp.y / tm;
    T anx = f.x / tm, any = f.y / tm;
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
    if (do_reset) {
      V2<T> z;
      z.x = T(0);
      z.y = T(0);
      A.force[i] = z;
    }
    if (model) {
      const double I = A.inertia[i];
      double ao = A.torque_prev[i] / I;
      double an = A.torque[i] / I;
      A.torque_prev[i] = A.torque[i];
      double w = A.omega[i] + (ao + an) / 2.0 * dt;
      double ph = A.phi[i] + (w * dt + an / 2.0 * (dt * dt));
      ph = wrap_to_pi(ph);
      A.omega[i] = w;
      A.phi[i] = ph;
      const double rts = A.r_ts[i];
      double ox = sin(ph) * rts, oy = (-cos(ph)) * rts;
      double2 l, r;
      l.x = x.x - ox;
      l.y = x.y - oy;
      r.x = x.x + ox;
      r.y = x.y + oy;
      A.ls[i] = l;
      A.rs[i] = r;
      if (do_reset) A.torque[i] = 0.0;
    }
    if (fused) {
      if (!(isfinite(x.x) && isfinite(x.y))) v[R_NONFINITE] = 1ull;
      unsigned long long bx = ord_bits(x.x), by = ord_bits(x.y);
      v[R_MINX] = bx;
      v[R_MINY] = by;
      v[R_MAXX] = bx;
      v[R_MAXY] = by;
      double l = dhypot((double)vel.x, (double)vel.y);
      if (l > vmax) vmax = l;
      v[R_V0MAX] = ord_bits(A.v0[i]);
    }
  }
  if (fused) {
    v[R_VMAX] = (unsigned long long)__double_as_longlong(vmax);
    if (block_reduce_commit(st, v, 1, partials) && threadIdx.x == 0) {
      st->dt = dt;
      st->time_tot += dt;
      st->step += 1;
      st->C_clear = 0;
    }
  } else if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->dt = dt;
    st->time_tot += dt;
  }
}

// ---------------------------------------------------------------------------
// Host helpers
// ---------------------------------------------------------------------------

inline int grid_for(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

inline int check_launch() { return hipGetLastError() == hipSuccess ? CD_OK : CD_ERR_HIP; }

int check_agents(const cd_agents_t *a) {
  if (!a) return CD_ERR_ARG;
  if (a->model != CD_CIRCULAR && a->model != CD_THREE_CIRCLE) return CD_ERR_TYPE;
  if (a->real != CD_F64 && a->real != CD_F32) return CD_ERR_ARG;
  if (a->n < 0 || a->n > (int64_t)0x7fffffff) return CD_ERR_ARG;
  if (a->model == CD_THREE_CIRCLE && a->real != CD_F64) return CD_ERR_UNSUPPORTED;
  if (a->n == 0) return CD_OK;
  if (!a->position || !a->velocity || !a->target_direction || !a->force || !a->force_prev ||
      !a->radius || !a->mass || !a->target_velocity || !a->tau_adj || !a->k_soc || !a->tau_0 ||
      !a->mu || !a->kappa || !a->damping || !a->std_rand_force || !a->target)
    return CD_ERR_ARG;
  if (a->model == CD_THREE_CIRCLE &&
      (!a->position_ls || !a->position_rs || !a->r_t || !a->r_s || !a->r_ts ||
       !a->inertia_rot || !a->target_angular_velocity || !a->orientation ||
       !a->angular_velocity || !a->target_orientation || !a->torque || !a->torque_prev ||
       !a->tau_rot || !a->std_rand_torque))
    return CD_ERR_ARG;
  return CD_OK;
}

int check_ws(const cd_agents_t *a, void *ws, int64_t ws_bytes, int64_t cap) {
  if (!ws || cap <= 0 || cap > (int64_t)0x3fffffff) return CD_ERR_ARG;
  Layout L = make_layout(a->n, cap, a->model);
  if (ws_bytes < L.total) return CD_ERR_WORKSPACE;
  return CD_OK;
}

constexpr int kBS = 256;
int g_pair_block = 128;  // threads per block of the pair kernels (autotuned)
int g_reduce_blocks = 512;

template <typename T>
int launch_reduce(const Agents<T> &A, const Work &W, hipStream_t s) {
  int blocks = grid_for(A.n, kBS);
  if (blocks > g_reduce_blocks) blocks = g_reduce_blocks;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_reduce<T>, dim3(blocks), dim3(kBS), 0, s, A, W.st, W.count,
                     W.partials);
  return check_launch();
}

// Optional per-kernel timing: when `ev` is not null, ev[k] is recorded
// before kernel k and ev[k+1] after it (events on the launch stream).
inline void mark(hipEvent_t *ev, int k, hipStream_t s) {
  if (ev) (void)hipEventRecord(ev[k], s);
}

template <typename T>
int launch_binning(const Agents<T> &A, const Work &W, double cell, hipStream_t s,
                   hipEvent_t *ev = nullptr) {
  const int n = A.n;
  mark(ev, 0, s);
  hipLaunchKernelGGL(k_bin_count, dim3(grid_for(n, kBS)), dim3(kBS), 0, s,
                     (const double2 *)A.pos, n, cell, W.st, W.count, W.cell_of, W.rank, W.cap);
  mark(ev, 1, s);
  hipLaunchKernelGGL(k_bin_scan, dim3(W.ntiles), dim3(kScanThreads), 0, s, W.st,
                     (const int *)W.count, W.offset, W.tiles);
  mark(ev, 2, s);
  hipLaunchKernelGGL(k_bin_scatter, dim3(grid_for(n, kBS)), dim3(kBS), 0, s, n,
                     (const DevState *)W.st, (const int *)W.cell_of, (const int *)W.rank,
                     (const int *)W.offset, W.pidx);
  mark(ev, 3, s);
  // one thread per cell of the capacity (the live cell count is only known
  // on the device; threads past it return at once)
  int64_t cells = W.cap;
  hipLaunchKernelGGL(k_bin_sort_gather<T>, dim3(grid_for(cells, kBS)), dim3(kBS), 0, s, A,
                     (const DevState *)W.st, (const int *)W.offset, W.pidx, W.scell, W.spos,
                     (V2<T> *)W.svel, W.sr, W.sls, W.srs, W.sphi, W.srt, W.srs_r, W.srts);
  mark(ev, 4, s);
  return check_launch();
}

template <typename T>
int launch_forces(const cd_agents_t *a, const Work &W, DevState *st, const FusedArgs &fa,
                  bool fused, hipStream_t s) {
  Agents<T> A = typed<T>(a);
  const Slots S = slots_of(W);
  const int threads = 4 * (int)a->n;
  const dim3 grid(grid_for(threads, g_pair_block)), block(g_pair_block);
  if (a->model == CD_THREE_CIRCLE) {
    if (fused)
      hipLaunchKernelGGL((k_forces<T, 1, true>), grid, block, 0, s, A, st, fa, S);
    else
      hipLaunchKernelGGL((k_forces<T, 1, false>), grid, block, 0, s, A, st, fa, S);
  } else {
    if (fused)
      hipLaunchKernelGGL((k_forces<T, 0, true>), grid, block, 0, s, A, st, fa, S);
    else
      hipLaunchKernelGGL((k_forces<T, 0, false>), grid, block, 0, s, A, st, fa, S);
  }
  return check_launch();
}

template <typename T>
int do_agent_agent(const cd_agents_t *a, double cell, void *ws, int64_t cap, hipStream_t s) {
  Agents<T> A = typed<T>(a);
  Layout L = make_layout(a->n, cap, a->model);
  Work W = carve(ws, L, cap);
  int rc = launch_reduce(A, W, s);
  if (rc) return rc;
  rc = launch_binning(A, W, cell, s);
  if (rc) return rc;
  FusedArgs fa;
  memset(&fa, 0, sizeof(fa));
  return launch_forces<T>(a, W, W.st, fa, false, s);
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

int64_t cd_workspace_bytes(int64_t n, int64_t cell_capacity, int32_t model, cd_layout_t *layout) {
  if (n < 0 || cell_capacity <= 0) return -1;
  Layout L = make_layout(n, cell_capacity, model);
  if (layout) {
    layout->points_indices = L.pidx;
    layout->cells_count = L.count;
    layout->cells_offset = L.offset;
    layout->cell_of = L.cell_of;
    layout->total_bytes = L.total;
  }
  return L.total;
}

int cd_workspace_init(void *ws, int64_t ws_bytes, int64_t n, int64_t cap, int32_t model,
                      cd_stream_t stream) {
  if (!ws || n < 0 || cap <= 0) return CD_ERR_ARG;
  Layout L = make_layout(n, cap, model);
  if (ws_bytes < L.total) return CD_ERR_WORKSPACE;
  Work W = carve(ws, L, cap);
  hipLaunchKernelGGL(k_ws_init, dim3(grid_for(cap + 1, kBS)), dim3(kBS), 0, (hipStream_t)stream,
                     W.st, W.count, (int)cap, W.tiles, W.ntiles);
  return check_launch();
}

int cd_read_state(const void *ws, cd_state_t *out, cd_stream_t stream) {
  if (!ws || !out) return CD_ERR_ARG;
  DevState h;
  if (hipMemcpyAsync(&h, ws, sizeof(DevState), hipMemcpyDeviceToHost, (hipStream_t)stream) !=
      hipSuccess)
    return CD_ERR_HIP;
  if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) return CD_ERR_HIP;
  out->dt = h.dt;
  out->time_tot = h.time_tot;
  out->origin[0] = h.origin[0];
  out->origin[1] = h.origin[1];
  out->shape[0] = h.nx;
  out->shape[1] = h.ny;
  out->status = h.status;
  out->step = (int64_t)h.step;
  double vm;
  memcpy(&vm, &h.red[0][R_VMAX], 8);
  out->v_max = vm;
  unsigned long long b = h.red[0][R_V0MAX];
  b = (b & 0x8000000000000000ull) ? (b & 0x7fffffffffffffffull) : ~b;
  memcpy(&vm, &b, 8);
  out->v0_max = vm;
  return CD_OK;
}

}  // extern "C"

namespace {
__global__ void k_clear_status(DevState *st) { st->status = 0; }
__global__ void k_philox(const uint32_t *in, uint32_t *out) {
  uint32_t c[4] = {in[0], in[1], in[2], in[3]};
  philox4x32(c, in[4], in[5]);
  for (int k = 0; k < 4; ++k) out[k] = c[k];
}
}  // namespace

extern "C" {

int cd_clear_status(void *ws, cd_stream_t stream) {
  if (!ws) return CD_ERR_ARG;
  hipLaunchKernelGGL(k_clear_status, dim3(1), dim3(1), 0, (hipStream_t)stream, (DevState *)ws);
  return check_launch();
}

int cd_add_to_cells(const cd_agents_t *a, double cell_size, void *ws, int64_t ws_bytes,
                    int64_t cap, cd_stream_t stream) {
  int rc = check_agents(a);
  if (rc) return rc;
  if (!(cell_size > 0.0)) return CD_ERR_ARG;
  if ((rc = check_ws(a, ws, ws_bytes, cap))) return rc;
  if (a->n == 0) return CD_OK;
  hipStream_t s = (hipStream_t)stream;
  Layout L = make_layout(a->n, cap, a->model);
  Work W = carve(ws, L, cap);
  if (a->real == CD_F64) {
    Agents<double> A = typed<double>(a);
    if ((rc = launch_reduce(A, W, s))) return rc;
    return launch_binning(A, W, cell_size, s);
  }
  Agents<float> A = typed<float>(a);
  if ((rc = launch_reduce(A, W, s))) return rc;
  return launch_binning(A, W, cell_size, s);
}

int cd_agent_agent_block_list(const cd_agents_t *a, double cell_size, void *ws, int64_t ws_bytes,
                              int64_t cap, cd_stream_t stream) {
  int rc = check_agents(a);
  if (rc) return rc;
  if (!(cell_size > 0.0)) return CD_ERR_ARG;
  if ((rc = check_ws(a, ws, ws_bytes, cap))) return rc;
  if (a->n == 0) return CD_OK;
  if (a->real == CD_F64)
    return do_agent_agent<double>(a, cell_size, ws, cap, (hipStream_t)stream);
  return do_agent_agent<float>(a, cell_size, ws, cap, (hipStream_t)stream);
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

int cd_velocity_verlet_integrator(const cd_agents_t *a, double dt_min, double dt_max, void *ws,
                                  int64_t ws_bytes, cd_stream_t stream) {
  int rc = check_agents(a);
  if (rc) return rc;
  const int64_t need = align_up(4096) + align_up(partials_bytes(a->n));
  if (!ws || ws_bytes < need) return CD_ERR_WORKSPACE;
  if (!(dt_min >= 0.0) || !(dt_max >= 0.0)) return CD_ERR_ARG;
  if (a->n == 0) return CD_OK;
  hipStream_t s = (hipStream_t)stream;
  DevState *st = (DevState *)ws;
  unsigned long long *partials = (unsigned long long *)((char *)ws + align_up(4096));
  const int rblocks = std::min(grid_for(a->n, kBS), g_reduce_blocks);
  if (a->real == CD_F64) {
    Agents<double> A = typed<double>(a);
    // count clearing is skipped (null counts): the integrator owns no binning
    hipLaunchKernelGGL(k_reduce<double>, dim3(rblocks), dim3(kBS), 0, s, A, st, (int *)nullptr,
                       partials);
    hipLaunchKernelGGL(k_integrate<double>, dim3(grid_for(A.n, kBS)), dim3(kBS), 0, s, A,
                       (int)a->model, st, dt_min, dt_max, 0, 0, (int *)nullptr, partials);
  } else {
    Agents<float> A = typed<float>(a);
    hipLaunchKernelGGL(k_reduce<float>, dim3(rblocks), dim3(kBS), 0, s, A, st, (int *)nullptr,
                       partials);
    hipLaunchKernelGGL(k_integrate<float>, dim3(grid_for(A.n, kBS)), dim3(kBS), 0, s, A,
                       (int)a->model, st, dt_min, dt_max, 0, 0, (int *)nullptr, partials);
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

}  // extern "C"

namespace {

int check_step(const cd_agents_t *a, const cd_step_params_t *p, void *ws, int64_t ws_bytes) {
  int rc = check_agents(a);
  if (rc) return rc;
  if (!p) return CD_ERR_ARG;
  if ((p->flags & CD_USE_AGENT_AGENT) && !(p->cell_size > 0.0)) return CD_ERR_ARG;
  if ((p->flags & CD_USE_OBSTACLE) && (p->n_segments < 0 || (p->n_segments > 0 && !p->segments)))
    return CD_ERR_ARG;
  if ((p->flags & CD_USE_NAV) &&
      (!p->nav_maps || p->nav_rows < 2 || p->nav_cols < 2 || !(p->nav_step > 0.0) ||
       (p->nav_mode != 0 && p->nav_mode != 1)))
    return CD_ERR_ARG;
  return check_ws(a, ws, ws_bytes, p->cell_capacity);
}

template <typename T>
int step_impl(const cd_agents_t *a, const cd_step_params_t *p, void *ws, hipStream_t s,
              hipEvent_t *ev = nullptr) {
  Agents<T> A = typed<T>(a);
  Layout L = make_layout(a->n, p->cell_capacity, a->model);
  Work W = carve(ws, L, p->cell_capacity);
  int rc;
  const bool pairs = (p->flags & CD_USE_AGENT_AGENT) != 0;
  if (pairs) {
    if ((rc = launch_binning(A, W, p->cell_size, s, ev))) return rc;
  } else {
    for (int k = 0; k < 5; ++k) mark(ev, k, s);
  }
  FusedArgs fa;
  fa.flags = p->flags;
  if (!pairs) fa.flags &= ~CD_USE_AGENT_AGENT;
  fa.nav_mode = p->nav_mode;
  fa.seed = p->seed;
  fa.maps = p->nav_maps;
  fa.nav_T = p->nav_targets;
  fa.nav_rows = p->nav_rows;
  fa.nav_cols = p->nav_cols;
  fa.nav_minx = p->nav_minx;
  fa.nav_miny = p->nav_miny;
  fa.nav_step = p->nav_step;
  fa.segs = (const double4 *)p->segments;
  fa.W = (p->flags & CD_USE_OBSTACLE) ? (int)p->n_segments : 0;
  if ((rc = launch_forces<T>(a, W, W.st, fa, true, s))) return rc;
  mark(ev, 5, s);
  if ((rc = check_launch())) return rc;
  hipLaunchKernelGGL(k_integrate<T>, dim3(grid_for(A.n, kBS)), dim3(kBS), 0, s, A,
                     (int)a->model, W.st, p->dt_min, p->dt_max, 1,
                     (p->flags & CD_USE_RESET) ? 1 : 0, W.count, W.partials);
  mark(ev, 6, s);
  return check_launch();
}

__global__ void k_set_step(DevState *st, unsigned long long step) { st->step = step; }

}  // namespace

extern "C" {

int cd_step_prepare(const cd_agents_t *a, const cd_step_params_t *p, void *ws, int64_t ws_bytes,
                    uint64_t first_step, cd_stream_t stream) {
  int rc = check_step(a, p, ws, ws_bytes);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  Layout L = make_layout(a->n, p->cell_capacity, a->model);
  Work W = carve(ws, L, p->cell_capacity);
  hipLaunchKernelGGL(k_set_step, dim3(1), dim3(1), 0, s, W.st, (unsigned long long)first_step);
  if ((rc = check_launch())) return rc;
  if (a->n == 0) return CD_OK;
  if (a->real == CD_F64) return launch_reduce(typed<double>(a), W, s);
  return launch_reduce(typed<float>(a), W, s);
}

int cd_step(const cd_agents_t *a, const cd_step_params_t *p, void *ws, int64_t ws_bytes,
            cd_stream_t stream) {
  int rc = check_step(a, p, ws, ws_bytes);
  if (rc) return rc;
  if (a->n == 0) return CD_OK;
  if (a->real == CD_F64) return step_impl<double>(a, p, ws, (hipStream_t)stream);
  return step_impl<float>(a, p, ws, (hipStream_t)stream);
}

int cd_step_timed(const cd_agents_t *a, const cd_step_params_t *p, void *ws, int64_t ws_bytes,
                  int64_t steps, double *kernel_ms, cd_stream_t stream) {
  int rc = check_step(a, p, ws, ws_bytes);
  if (rc) return rc;
  if (!kernel_ms || steps < 0) return CD_ERR_ARG;
  for (int k = 0; k < 6; ++k) kernel_ms[k] = 0.0;
  if (a->n == 0 || steps == 0) return CD_OK;
  hipStream_t s = (hipStream_t)stream;
  hipEvent_t ev[7];
  for (int k = 0; k < 7; ++k)
    if (hipEventCreate(&ev[k]) != hipSuccess) return CD_ERR_HIP;
  for (int64_t it = 0; it < steps && rc == CD_OK; ++it) {
    rc = a->real == CD_F64 ? step_impl<double>(a, p, ws, s, ev) : step_impl<float>(a, p, ws, s, ev);
    if (rc) break;
    if (hipEventSynchronize(ev[6]) != hipSuccess) {
      rc = CD_ERR_HIP;
      break;
    }
    for (int k = 0; k < 6; ++k) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, ev[k], ev[k + 1]) != hipSuccess) rc = CD_ERR_HIP;
      kernel_ms[k] += ms;
    }
  }
  for (int k = 0; k < 7; ++k) (void)hipEventDestroy(ev[k]);
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
  hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
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

int cd_philox4x32(const uint32_t *in, uint32_t *out, cd_stream_t stream) {
  if (!in || !out) return CD_ERR_ARG;
  hipLaunchKernelGGL(k_philox, dim3(1), dim3(1), 0, (hipStream_t)stream, in, out);
  return check_launch();
}

const char *cd_version(void) { return "cdamd 0.1 gfx950"; }

}  // extern "C"
