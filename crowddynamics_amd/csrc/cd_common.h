// Shared part of the step kernels (cd_kernels.hip) and the adjoint
// (cd_adjoint.hip): workspace layout, device state, the scenario view of a
// launch, reductions, binning (K_a, K_b) and argument checks.
//
// Data layout in HBM: one contiguous array per agent field (vectors
// interleaved [n][2]), plus a caller-owned workspace holding the device state
// scalars, the scenario table and per-scenario state of a batch, per-block
// reduction partials, per-cell agent slot lists with gathered neighbour data,
// and the block-list arrays of cd_add_to_cells.
//
// Binning contract (the reference binds cell_lists, an un-vendored
// dependency, at crowddynamics/core/interactions.py:30,192-195):
//   origin o = per-axis minimum of the positions, l = trunc((p - o) / c),
//   shape = max l + 1, flat cell = l_x * ny + l_y, agents ordered stably by
//   flat cell (ties in agent index order), and every unordered pair of
//   agents in the same or adjacent cells (|dl_x| <= 1 and |dl_y| <= 1)
//   interacts once.

#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "../../include/cdamd.h"
#include "cd_device.h"

namespace {

using namespace cd;


// ---------------------------------------------------------------------------
// Workspace
// ---------------------------------------------------------------------------

enum RedField { R_MINX = 0, R_MINY, R_MAXX, R_MAXY, R_VMAX, R_V0MAX, R_NONFINITE, R_COUNT };

// Global device state.
struct DevState {
  long long status;
  unsigned long long step;
  unsigned int scan_epoch, scan_ticket;
};

// Scenario table entry (device).
struct ScnDesc {
  int a0, a1;    // agents [a0, a1)
  int c0, ccap;  // grid cells [c0, c0 + ccap)
  int w0, w1;    // wall segments [w0, w1)
  int f0, nf;    // blocks (and partials) [f0, f0 + nf) of the per-scenario kernels
  int nt;        // blocks of the tile force kernel
  unsigned long long seed;
  double dt_min, dt_max;
};

// Per-scenario state (device).  The fused step keeps two cell-count
// buffers and two reduction slots, selected by the parity of the step: the
// step's first kernel fills buffer p and clears buffer 1 - p; its force
// kernel merges the reduction of the new state into slot 1 - p, which the
// next step reads.  A slot is kAccSub sub-slots (Work::acc), so that the
// merging blocks spread their atomics.  The operator paths use buffer 0
// only.
struct ScnState {
  unsigned long long red[8];  // last reduced values (R_*), ord-bit form except R_VMAX
  double dt, time_tot;
  double origin[2];
  int nx, ny;
  int valid;       // grid of the last binning fits the scenario's cells
  int C;           // cells of the last binning
  int C_clear[2];  // cells of count buffer b that its next clearing pass must zero
  int oob[2];      // an agent fell outside the grid (stale bounding box), by buffer
  int par;         // count buffer and reduction slot of the current fused step
};

constexpr int kBS = 256;              // threads per block of per-agent kernels
constexpr int kAccSub = 128;          // sub-slots of a scenario's reduction slot
constexpr int kMaxPartsPerScn = 128;  // blocks of the partial-producing kernels per scenario
constexpr int kSingleParts = 512;     // the same for a single (unbatched) scenario
constexpr int kAgentsPerForceBlock = 64;  // default agents per force block (256 threads)
constexpr int kMinForceAgents = 4;  // agents per force block: 64 threads / 16 lanes
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

constexpr int64_t kAlign = 256;
inline int64_t align_up(int64_t x) { return (x + kAlign - 1) / kAlign * kAlign; }
inline int grid_for(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

inline int64_t max_blocks_f(int64_t n, int64_t S) {
  return std::max<int64_t>(kSingleParts, n / kBS + S + 1);
}
inline int64_t max_blocks_e(int64_t n, int64_t S) { return n / kMinForceAgents + S + 1; }
// lanes per agent of the agent-parallel force kernel (2, 4, 8, 16, 32 or 64; 0: 4)
inline int agent_lanes(const cd_workspace_t *w) { return w->agent_lanes > 0 ? (int)w->agent_lanes : 4; }

struct Layout {
  int64_t state, scn_desc, scn_state, acc, map_f, map_e, map_t, partials, tiles;
  int64_t cell_of, rank, pidx, count, offset;
  int64_t ell_ids, e_pos, e_vel, e_r, e_ls, e_rs, e_phi, e_rt, e_rs_r, e_rts;
  int64_t snap_pos, snap_vel, snap_ls, snap_rs, snap_off;
  int64_t order, order_off;
  int64_t rec;
  int64_t total;
};

Layout make_layout(const cd_workspace_t *w) {
  const int64_t n = w->n, cap = w->cell_capacity, K = w->cell_slots;
  const int64_t S = std::max<int64_t>(1, w->max_scenarios);
  Layout L;
  int64_t o = 0;
  auto take = [&](int64_t bytes) {
    int64_t r = o;
    o = align_up(o + bytes);
    return r;
  };
  L.state = take(256);
  L.scn_desc = take(sizeof(ScnDesc) * S);
  L.scn_state = take(sizeof(ScnState) * S);
  L.acc = take(8 * 8 * kAccSub * 2 * S);  // [scenario][slot][sub-slot][8]
  L.map_f = take(8 * max_blocks_f(n, S));
  L.map_e = take(8 * max_blocks_e(n, S));
  L.map_t = take(8 * (cap + S + 1));  // at most one tile block per cell, plus one per scenario
  L.partials = take(8 * 8 * max_blocks_f(n, S));
  L.tiles = take(8 * ((cap + kScanTile - 1) / kScanTile + 1));
  L.cell_of = take(4 * n);
  L.rank = take(4 * n);
  L.pidx = take(4 * n);
  L.count = take(4 * 2 * cap);  // two buffers
  L.offset = take(4 * (cap + 1));
  const int64_t slots = cap * K;
  L.ell_ids = take(4 * slots);
  L.e_pos = take(16 * slots);
  L.e_vel = take(16 * slots);
  L.e_r = take(8 * slots);
  if (w->model == CD_THREE_CIRCLE) {
    L.e_ls = take(16 * slots);
    L.e_rs = take(16 * slots);
    L.e_phi = take(8 * slots);
    L.e_rt = take(8 * slots);
    L.e_rs_r = take(8 * slots);
    L.e_rts = take(8 * slots);
  } else {
    L.e_ls = L.e_rs = L.e_phi = L.e_rt = L.e_rs_r = L.e_rts = -1;
  }
  // the state a fused step starts from, read as neighbour data while the
  // force kernel integrates
  L.snap_pos = take(16 * n);
  L.snap_vel = take(16 * n);
  if (w->model == CD_THREE_CIRCLE) {
    L.snap_ls = take(16 * n);
    L.snap_rs = take(16 * n);
    L.snap_off = take(16 * n);  // shoulder offsets
  } else {
    L.snap_ls = L.snap_rs = L.snap_off = -1;
  }
  // processing order of the agent-parallel force kernel and the cell
  // offsets it is built from
  L.order = take(4 * n);
  L.order_off = take(4 * cap);
  // circular agents: the fused step's neighbour records, one per slot
  // (NbrRec: 48 bytes in fp64, 32 in fp32)
  L.rec = (w->model == CD_CIRCULAR) ? take(48 * slots) : -1;
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

// Device views of the workspace arrays (passed by value to kernels).
struct Work {
  DevState *st;
  unsigned long long *acc;  // reduction slots of the fused step
  unsigned long long *partials, *tiles;
  int *cell_of, *rank, *pidx, *count, *offset;
  int *ell_ids;
  double2 *e_pos;
  void *e_vel;
  double *e_r;
  double2 *e_ls, *e_rs;
  double *e_phi, *e_rt, *e_rs_r, *e_rts;
  double2 *snap_pos;
  void *snap_vel;
  double2 *snap_ls, *snap_rs;
  double2 *snap_off;
  int *order, *order_off;  // force kernel's processing order (cell order when order_every > 0)
  void *rec;               // circular agents: NbrRec of each slot, written by the fused step's K_a
  int cap, K, ntiles;
  int order_every;
};

// Scalars of an adjoint or tangent step with adaptive dt, in the workspace's
// state region after DevState: the step's dt and the maxima it was formed
// from (core/integrator.py:29-97), whether dt lies between its clamps (its
// derivative is then not zero), the lowest agent index attaining each
// maximum, the exact sum of dL/d(dt) over the agents (adjoint, as a 128-bit
// fixed-point pair) and the tangent of dt (tangent step).
struct AdjScal {
  double dt, vmax, v0max, ddt;
  int inside, arg_v, arg_v0, bad;
  unsigned long long ldt_lo;
  long long ldt_hi;
};
constexpr int kAdjScalOffset = 64;
static_assert(sizeof(DevState) <= kAdjScalOffset && kAdjScalOffset + sizeof(AdjScal) <= 256,
              "the adjoint scalars share the 256-byte state region");

// Scenario view of a launch: the table (null for a single scenario, which is
// then `one`), the per-scenario states and the block maps (null: block b of a
// single scenario is its b-th block).
struct Batch {
  const ScnDesc *desc;
  ScnState *sst;
  const int2 *map_f, *map_e, *map_t;
  ScnDesc one;
};

Work carve(const cd_workspace_t *w) {
  const Layout L = make_layout(w);
  char *b = (char *)w->ptr;
  Work W;
  W.st = (DevState *)(b + L.state);
  W.acc = (unsigned long long *)(b + L.acc);
  W.partials = (unsigned long long *)(b + L.partials);
  W.tiles = (unsigned long long *)(b + L.tiles);
  W.cell_of = (int *)(b + L.cell_of);
  W.rank = (int *)(b + L.rank);
  W.pidx = (int *)(b + L.pidx);
  W.count = (int *)(b + L.count);
  W.offset = (int *)(b + L.offset);
  W.ell_ids = (int *)(b + L.ell_ids);
  W.e_pos = (double2 *)(b + L.e_pos);
  W.e_vel = (void *)(b + L.e_vel);
  W.e_r = (double *)(b + L.e_r);
  W.e_ls = L.e_ls >= 0 ? (double2 *)(b + L.e_ls) : nullptr;
  W.e_rs = L.e_rs >= 0 ? (double2 *)(b + L.e_rs) : nullptr;
  W.e_phi = L.e_phi >= 0 ? (double *)(b + L.e_phi) : nullptr;
  W.e_rt = L.e_rt >= 0 ? (double *)(b + L.e_rt) : nullptr;
  W.e_rs_r = L.e_rs_r >= 0 ? (double *)(b + L.e_rs_r) : nullptr;
  W.e_rts = L.e_rts >= 0 ? (double *)(b + L.e_rts) : nullptr;
  W.snap_pos = (double2 *)(b + L.snap_pos);
  W.snap_vel = (void *)(b + L.snap_vel);
  W.snap_ls = L.snap_ls >= 0 ? (double2 *)(b + L.snap_ls) : nullptr;
  W.snap_rs = L.snap_rs >= 0 ? (double2 *)(b + L.snap_rs) : nullptr;
  W.snap_off = L.snap_off >= 0 ? (double2 *)(b + L.snap_off) : nullptr;
  W.order = (int *)(b + L.order);
  W.rec = L.rec >= 0 ? (void *)(b + L.rec) : nullptr;
  W.order_off = (int *)(b + L.order_off);
  // only the agent-parallel force kernel follows the order
  W.order_every = (w->tile_x > 0 && w->tile_y > 0) ? 0 : (int)w->order_every;
  W.cap = (int)w->cell_capacity;
  W.K = (int)w->cell_slots;
  W.ntiles = (int)((w->cell_capacity + kScanTile - 1) / kScanTile);
  return W;
}

__host__ __device__ __forceinline__ AdjScal *adj_scal(const Work &W) {
  return (AdjScal *)((char *)W.st + kAdjScalOffset);
}

inline int single_parts(int64_t n) {
  return std::max(1, std::min(grid_for(n, kBS), kSingleParts));
}

// Blocks of the tiled force kernel for a scenario of `cells` grid cells at
// most (blocks stride over the tiles of the actual grid); one when the
// workspace selects the direct kernel (tile_x or tile_y 0).
constexpr int kMaxTileBlocks = 4096;
inline int tile_blocks(int64_t cells, const cd_workspace_t *w) {
  if (w->tile_x <= 0 || w->tile_y <= 0) return 1;
  const int64_t per = (int64_t)w->tile_x * w->tile_y;
  return (int)std::max<int64_t>(1, std::min<int64_t>((cells + per - 1) / per, kMaxTileBlocks));
}

// Batch view of a single scenario covering every agent, cell and wall.
Batch single_batch(const cd_workspace_t *w, int64_t n_segments, uint64_t seed, double dt_min,
                   double dt_max) {
  const Layout L = make_layout(w);
  Batch B;
  B.desc = nullptr;
  B.sst = (ScnState *)((char *)w->ptr + L.scn_state);
  B.map_f = nullptr;
  B.map_e = nullptr;
  B.map_t = nullptr;
  B.one.a0 = 0;
  B.one.a1 = (int)w->n;
  B.one.c0 = 0;
  B.one.ccap = (int)w->cell_capacity;
  B.one.w0 = 0;
  B.one.w1 = (int)n_segments;
  B.one.f0 = 0;
  B.one.nf = single_parts(w->n);
  B.one.nt = tile_blocks(w->cell_capacity, w);
  B.one.seed = seed;
  B.one.dt_min = dt_min;
  B.one.dt_max = dt_max;
  return B;
}

Batch batch_of(const cd_workspace_t *w) {
  const Layout L = make_layout(w);
  char *b = (char *)w->ptr;
  Batch B;
  memset(&B, 0, sizeof(B));
  B.desc = (const ScnDesc *)(b + L.scn_desc);
  B.sst = (ScnState *)(b + L.scn_state);
  B.map_f = (const int2 *)(b + L.map_f);
  B.map_e = (const int2 *)(b + L.map_e);
  B.map_t = (const int2 *)(b + L.map_t);
  return B;
}

__device__ __forceinline__ ScnDesc desc_of(const Batch &B, int s) {
  return B.desc ? B.desc[s] : B.one;
}
// (scenario, block index within the scenario) of this block
__device__ __forceinline__ int2 block_f(const Batch &B) {
  return B.map_f ? B.map_f[blockIdx.x] : make_int2(0, (int)blockIdx.x);
}
__device__ __forceinline__ int2 block_e(const Batch &B) {
  return B.map_e ? B.map_e[blockIdx.x] : make_int2(0, (int)blockIdx.x);
}
// The same for kernels instantiated for one scenario (ONE: single_batch,
// whose table and maps are null) or for a batch (batch_of, whose table and
// maps are all set): the descriptor comes straight from the kernel
// arguments or straight from the table.  A descriptor read that may come
// from either makes each of its scalar loads wait for the previous one.
template <bool ONE>
__device__ __forceinline__ ScnDesc desc_t(const Batch &B, int s) {
  if constexpr (ONE)
    return B.one;
  else
    return B.desc[s];
}
template <bool ONE>
__device__ __forceinline__ int2 block_f_t(const Batch &B) {
  if constexpr (ONE)
    return make_int2(0, (int)blockIdx.x);
  else
    return B.map_f[blockIdx.x];
}
template <bool ONE>
__device__ __forceinline__ int2 block_e_t(const Batch &B) {
  if constexpr (ONE)
    return make_int2(0, (int)blockIdx.x);
  else
    return B.map_e[blockIdx.x];
}

constexpr unsigned long long kOrdPosInf = 0xFFF0000000000000ull;  // ord_bits(+inf)
constexpr unsigned long long kOrdNegInf = 0x000FFFFFFFFFFFFFull;  // ord_bits(-inf)

__global__ void __launch_bounds__(256) k_ws_init(Work W, ScnState *sst, int S) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int stride = gridDim.x * blockDim.x;
  if (t == 0) memset(W.st, 0, sizeof(DevState));
  for (int s = t; s < S; s += stride) memset(&sst[s], 0, sizeof(ScnState));
  for (int c = t; c < 2 * W.cap; c += stride) W.count[c] = 0;
  for (int k = t; k < W.ntiles + 1; k += stride) W.tiles[k] = 0ull;
}

// ---------------------------------------------------------------------------
// Reductions (bounding box, max |v|, max v0, non-finite flag)
// ---------------------------------------------------------------------------

struct Red {
  unsigned long long v[R_COUNT];
  __device__ void init() {
    v[R_MINX] = kOrdPosInf;
    v[R_MINY] = kOrdPosInf;
    v[R_MAXX] = kOrdNegInf;
    v[R_MAXY] = kOrdNegInf;
    v[R_VMAX] = 0ull;  // bits of a non-negative double
    v[R_V0MAX] = kOrdNegInf;
    v[R_NONFINITE] = 0ull;
  }
  __device__ void add_agent(double2 p, double speed, double v0) {
    if (!(isfinite(p.x) && isfinite(p.y))) v[R_NONFINITE] = 1ull;
    const unsigned long long bx = ord_bits(p.x), by = ord_bits(p.y);
    v[R_MINX] = min(v[R_MINX], bx);
    v[R_MINY] = min(v[R_MINY], by);
    v[R_MAXX] = max(v[R_MAXX], bx);
    v[R_MAXY] = max(v[R_MAXY], by);
    // strict >, NaN never selected (core/integrator.py:81-85)
    if (speed > __longlong_as_double((long long)v[R_VMAX]))
      v[R_VMAX] = (unsigned long long)__double_as_longlong(speed);
    v[R_V0MAX] = max(v[R_V0MAX], ord_bits(v0));
  }
  __device__ void combine(const unsigned long long *u) {
    v[R_MINX] = min(v[R_MINX], u[R_MINX]);
    v[R_MINY] = min(v[R_MINY], u[R_MINY]);
    v[R_MAXX] = max(v[R_MAXX], u[R_MAXX]);
    v[R_MAXY] = max(v[R_MAXY], u[R_MAXY]);
    v[R_VMAX] = max(v[R_VMAX], u[R_VMAX]);
    v[R_V0MAX] = max(v[R_V0MAX], u[R_V0MAX]);
    v[R_NONFINITE] = max(v[R_NONFINITE], u[R_NONFINITE]);
  }
};

__device__ __forceinline__ unsigned long long shfl_xor64(unsigned long long v, int o) {
  return (unsigned long long)__shfl_xor((long long)v, o, 64);
}

// Block-wide reduction; the result is returned to every thread.
__device__ Red block_reduce(Red r) {
  __shared__ unsigned long long sh[R_COUNT][16];
  __shared__ unsigned long long res[R_COUNT];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    unsigned long long u[R_COUNT];
#pragma unroll
    for (int k = 0; k < R_COUNT; ++k) u[k] = shfl_xor64(r.v[k], o);
    r.combine(u);
  }
  __syncthreads();
  if (lane == 0)
    for (int k = 0; k < R_COUNT; ++k) sh[k][wid] = r.v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < nw; ++w) {
      unsigned long long u[R_COUNT];
      for (int k = 0; k < R_COUNT; ++k) u[k] = sh[k][w];
      r.combine(u);
    }
    for (int k = 0; k < R_COUNT; ++k) res[k] = r.v[k];
  }
  __syncthreads();
  for (int k = 0; k < R_COUNT; ++k) r.v[k] = res[k];
  return r;
}

// Producer side: each block stores its reduction as one partial (index =
// its block).  The next kernel on the stream combines a scenario's
// partials; kernel boundaries make them visible, so producers need neither
// fences nor atomics.
__device__ __forceinline__ void store_partial(const Work &W, const Red &r) {
  if (threadIdx.x == 0)
    for (int k = 0; k < R_COUNT; ++k) W.partials[blockIdx.x * 8 + k] = r.v[k];
}

// Consumer side: every block combines its scenario's partials.
__device__ Red reduce_partials(const Work &W, const ScnDesc &d) {
  Red r;
  r.init();
  for (int b = threadIdx.x; b < d.nf; b += blockDim.x) r.combine(&W.partials[(d.f0 + b) * 8]);
  return block_reduce(r);
}

// Merges a block's reduction into a scenario's slot: exact and independent
// of order (min / max of ordered bits, or of the non-negative speed bits).
__device__ __forceinline__ void acc_merge(unsigned long long *acc, const Red &r) {
  atomicMin(&acc[R_MINX], r.v[R_MINX]);
  atomicMin(&acc[R_MINY], r.v[R_MINY]);
  atomicMax(&acc[R_MAXX], r.v[R_MAXX]);
  atomicMax(&acc[R_MAXY], r.v[R_MAXY]);
  atomicMax(&acc[R_VMAX], r.v[R_VMAX]);
  atomicMax(&acc[R_V0MAX], r.v[R_V0MAX]);
  atomicMax(&acc[R_NONFINITE], r.v[R_NONFINITE]);
}

__device__ __forceinline__ void acc_reset(unsigned long long *acc) {
  Red r;
  r.init();
  for (int k = 0; k < R_COUNT; ++k) acc[k] = r.v[k];
}

// Sub-slot `sub` of reduction slot `slot` of scenario `scn`.
__device__ __forceinline__ unsigned long long *acc_at(const Work &W, int scn, int slot, int sub) {
  return W.acc + ((size_t)(scn * 2 + slot) * kAccSub + sub) * 8;
}

// Wave-wide reduction; every lane returns the wave's result.  `first`: the
// lanes hold the identity except every first-th one (a power of two), so
// the exchanges below that distance are skipped.
__device__ __forceinline__ Red wave_reduce(Red r, int first = 1) {
  for (int o = 32; o >= first; o >>= 1) {
    unsigned long long u[R_COUNT];
#pragma unroll
    for (int k = 0; k < R_COUNT; ++k) u[k] = shfl_xor64(r.v[k], o);
    r.combine(u);
  }
  return r;
}

// The reduction held by a scenario's slot, combined by each wave on its own
// (no block barrier); every thread returns it.
__device__ __forceinline__ Red acc_read(const Work &W, int scn, int slot) {
  Red r;
  r.init();
  for (int sub = threadIdx.x & 63; sub < kAccSub; sub += 64) r.combine(acc_at(W, scn, slot, sub));
  return wave_reduce(r);
}

__device__ __forceinline__ double red_vmax(const unsigned long long *v) {
  return __longlong_as_double((long long)v[R_VMAX]);
}

// adaptive_timestep (core/integrator.py:81-97) from reduced values.
__device__ __forceinline__ double adaptive_dt(const unsigned long long *v, double dt_min,
                                              double dt_max) {
  const double v_max = red_vmax(v);
  if (v_max == 0.0) return dt_max;
  const double v0max = from_ord_bits(v[R_V0MAX]);
  const double dx_max = 1.1 * v0max * dt_max;
  const double dt = dx_max / v_max;
  if (dt > dt_max) return dt_max;
  if (dt < dt_min) return dt_min;
  return dt;
}

// Partials of the current state (operators and the step's preparation);
// also zeroes the counts of the previous binnings (both buffers).  acc_slot
// >= 0: the reduction is also merged into the scenario's acc[acc_slot]
// (reset beforehand by k_scn_prepare).
template <typename T>
__global__ void __launch_bounds__(256) k_reduce(Agents<T> A, Work W, Batch B, int clear_counts,
                                                int acc_slot) {
  const int2 sk = block_f(B);
  const ScnDesc d = desc_of(B, sk.x);
  ScnState *ss = &B.sst[sk.x];
  const int t0 = sk.y * blockDim.x + threadIdx.x, stride = d.nf * blockDim.x;
  if (clear_counts) {
    for (int b = 0; b < 2; ++b) {
      int *count = W.count + (size_t)b * W.cap + d.c0;
      const int clear = ss->C_clear[b];
      for (int c = t0; c < clear; c += stride) count[c] = 0;
    }
  }
  Red r;
  r.init();
  for (int i = d.a0 + t0; i < d.a1; i += stride) {
    const V2<T> w = A.vel[i];
    r.add_agent(A.pos[i], dhypot((double)w.x, (double)w.y), A.v0[i]);
  }
  r = block_reduce(r);
  store_partial(W, r);
  if (acc_slot >= 0 && threadIdx.x == 0)
    acc_merge(acc_at(W, sk.x, acc_slot, blockIdx.x % kAccSub), r);
  if (sk.y == 0 && threadIdx.x == 0) ss->oob[0] = 0;
}

// Before the first fused step: reset reduction slot `slot` of every
// scenario (k_reduce then fills it) and the out-of-grid flags.  One thread
// per (scenario, sub-slot).
__global__ void k_scn_prepare(Work W, ScnState *sst, int S, int slot) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= S * kAccSub) return;
  const int s = t / kAccSub, sub = t % kAccSub;
  acc_reset(acc_at(W, s, slot, sub));
  if (sub == 0) {
    sst[s].oob[0] = 0;
    sst[s].oob[1] = 0;
  }
}

// Grid of the binning from a reduced bounding box (contract in the header).
struct Grid {
  double ox, oy;
  int nx, ny;
  bool ok;
};

__device__ __forceinline__ Grid grid_of(const Red &r, double cell, int cap) {
  Grid g;
  const double mnx = from_ord_bits(r.v[R_MINX]), mny = from_ord_bits(r.v[R_MINY]);
  const double mxx = from_ord_bits(r.v[R_MAXX]), mxy = from_ord_bits(r.v[R_MAXY]);
  g.ox = mnx;
  g.oy = mny;
  const double qx = (mxx - mnx) / cell, qy = (mxy - mny) / cell;
  g.ok = r.v[R_NONFINITE] == 0ull && qx >= 0.0 && qy >= 0.0 && qx < (double)cap &&
         qy < (double)cap;
  g.nx = g.ny = 0;
  if (g.ok) {
    const long long nx = (long long)qx + 1, ny = (long long)qy + 1;
    g.ok = nx * ny <= (long long)cap;
    g.nx = (int)nx;
    g.ny = (int)ny;
  }
  return g;
}

// Records a scenario's reduction and grid (one thread per scenario); `buf`:
// the count buffer the binning fills.
__device__ __forceinline__ void publish_grid(const Work &W, ScnState *ss, const Red &r,
                                             const Grid &g, bool binning, bool has_agents,
                                             int buf) {
  for (int k = 0; k < R_COUNT; ++k) ss->red[k] = r.v[k];
  if (r.v[R_NONFINITE])
    atomicOr((unsigned long long *)&W.st->status, (unsigned long long)CD_STATUS_NONFINITE);
  if (!binning) {
    ss->valid = 0;
    return;
  }
  ss->origin[0] = g.ox;
  ss->origin[1] = g.oy;
  ss->nx = g.nx;
  ss->ny = g.ny;
  ss->valid = g.ok ? 1 : 0;
  ss->C = g.ok ? g.nx * g.ny : 0;
  ss->C_clear[buf] = g.ok ? g.nx * g.ny : 0;
  if (!g.ok && has_agents)  // an empty scenario has no grid and no overflow
    atomicOr((unsigned long long *)&W.st->status, (unsigned long long)CD_STATUS_GRID_OVERFLOW);
}

// Local cell of a position, or -1 outside the grid.
__device__ __forceinline__ int cell_of_pos(double2 p, const Grid &g, double cell) {
  const double qx = (p.x - g.ox) / cell, qy = (p.y - g.oy) / cell;
  if (!(qx >= 0.0 && qy >= 0.0 && qx < (double)g.nx && qy < (double)g.ny)) return -1;
  return (int)qx * g.ny + (int)qy;
}

// ---------------------------------------------------------------------------
// K_a / K_b: per-cell slot lists
// ---------------------------------------------------------------------------

// Binning of a scenario's agents with the grid of its reduced partials:
// publishes the grid (one thread per scenario) and, when `binning`, inserts
// the agents into their cells' slot lists (slots beyond cell_slots are not
// stored; the cell count tells).  Cell indices are global: scenario base +
// local cell.  p0: the position of the block's first agent, loaded early.
__device__ __forceinline__ void bin_scenario(const double2 *pos, double2 p0, double cell,
                                             const Work &W, const ScnDesc &d, ScnState *ss,
                                             int2 sk, const Red &r, int binning) {
  const int stride = d.nf * blockDim.x;
  const int i0 = d.a0 + sk.y * blockDim.x + threadIdx.x;
  const Grid g = grid_of(r, cell, d.ccap);
  if (sk.y == 0 && threadIdx.x == 0) publish_grid(W, ss, r, g, binning != 0, d.a1 > d.a0, 0);
  if (!binning || !g.ok) return;
  for (int i = i0; i < d.a1; i += stride) {
    const int lc = cell_of_pos(i == i0 ? p0 : pos[i], g, cell);
    if (lc < 0) {  // positions changed after the bounding box was reduced
      atomicOr(&ss->oob[0], 1);
      W.cell_of[i] = -1;
      continue;
    }
    const int c = d.c0 + lc;
    W.cell_of[i] = c;
    const int k = atomicAdd(&W.count[c], 1);
    if (k < W.K)
      W.ell_ids[c * W.K + k] = i;
    else if (k == W.K)  // the cell's list is full: its members are found by scanning cell_of
      atomicOr((unsigned long long *)&W.st->status, (unsigned long long)CD_STATUS_SLOTS_FULL);
  }
}

// K_a of the operator paths: combine the scenario's partials, then
// bin_scenario.
__global__ void __launch_bounds__(256) k_bin_insert(const double2 *pos, double cell, Work W,
                                                    Batch B, int binning) {
  const int2 sk = block_f(B);
  const ScnDesc d = desc_of(B, sk.x);
  ScnState *ss = &B.sst[sk.x];
  const int i0 = d.a0 + sk.y * blockDim.x + threadIdx.x;
  const double2 p0 = (binning && i0 < d.a1) ? pos[i0] : make_double2(0.0, 0.0);
  const Red r = reduce_partials(W, d);
  bin_scenario(pos, p0, cell, W, d, ss, sk, r, binning);
}

// K_b: one wave per cell of the workspace grid.  Orders the cell's slot list
// by agent index and gathers the neighbour data in that order.  Up to 64
// agents, each lane takes one id and finds its rank among the cell's ids by
// shuffles (ids are distinct, so the ranks are a permutation); larger cells
// are ordered by the wave's first lane.  Cells with more agents than slots
// are flagged (the adjoint visits their members through the exact slow
// path).
template <typename T>
__device__ __forceinline__ void gather_slot(const Agents<T> &A, const Work &W, size_t e, int i) {
  V2<T> *e_vel = (V2<T> *)W.e_vel;
  W.e_pos[e] = A.pos[i];
  e_vel[e] = A.vel[i];
  W.e_r[e] = A.radius[i];
  if (W.e_ls) {
    W.e_ls[e] = A.ls[i];
    W.e_rs[e] = A.rs[i];
    W.e_phi[e] = A.phi[i];
    W.e_rt[e] = A.r_t[i];
    W.e_rs_r[e] = A.r_s[i];
    W.e_rts[e] = A.r_ts[i];
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_cell_gather(Agents<T> A, Work W) {
  const int lane = threadIdx.x & 63;
  const int c = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6);
  if (c >= W.cap) return;  // whole waves
  const int cnt = W.count[c];
  if (cnt == 0) return;
  if (cnt > W.K) {
    if (lane == 0)
      atomicOr((unsigned long long *)&W.st->status, (unsigned long long)CD_STATUS_SLOTS_FULL);
    return;
  }
  int *ids = W.ell_ids + (size_t)c * W.K;
  if (cnt <= 64) {
    const int id = lane < cnt ? ids[lane] : 0x7fffffff;
    int rank = 0;
    for (int t = 0; t < cnt; ++t) rank += __shfl(id, t, 64) < id ? 1 : 0;
    if (lane < cnt) {
      ids[rank] = id;
      gather_slot(A, W, (size_t)c * W.K + rank, id);
    }
    return;
  }
  if (lane != 0) return;
  for (int s = 1; s < cnt; ++s) {  // insertion sort of a crowded cell
    const int key = ids[s];
    int t = s - 1;
    while (t >= 0 && ids[t] > key) {
      ids[t + 1] = ids[t];
      --t;
    }
    ids[t + 1] = key;
  }
  for (int s = 0; s < cnt; ++s) gather_slot(A, W, (size_t)c * W.K + s, ids[s]);
}

// ---------------------------------------------------------------------------
// Host helpers
// ---------------------------------------------------------------------------

inline int check_launch() { return hipGetLastError() == hipSuccess ? CD_OK : CD_ERR_HIP; }

int check_agents(const cd_agents_t *a) {
  if (!a) return CD_ERR_ARG;
  if (a->model != CD_CIRCULAR && a->model != CD_THREE_CIRCLE) return CD_ERR_TYPE;
  if (a->real != CD_F64 && a->real != CD_F32) return CD_ERR_ARG;
  if (a->n < 0 || a->n > (int64_t)0x1fffffff) return CD_ERR_ARG;
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

bool ws_sizes_ok(const cd_workspace_t *w) {
  return w && w->n >= 0 && w->n <= (int64_t)0x1fffffff && w->cell_capacity > 0 &&
         w->cell_capacity <= (int64_t)0x1fffffff && w->cell_slots > 0 &&
         w->cell_slots <= 4096 && w->cell_capacity * w->cell_slots <= ((int64_t)1 << 31) &&
         w->max_scenarios >= 0 && w->max_scenarios <= (1 << 20) &&
         (w->force_threads == 0 || w->force_threads == 64 || w->force_threads == 128 ||
          w->force_threads == 256) &&
         w->tile_x >= 0 && w->tile_x <= 16 && w->tile_y >= 0 && w->tile_y <= 16 &&
         (w->agent_lanes == 0 || w->agent_lanes == 2 || w->agent_lanes == 4 || w->agent_lanes == 8 ||
          w->agent_lanes == 16 || w->agent_lanes == 32 || w->agent_lanes == 64) &&
         w->order_every >= 0 &&
         (w->model == CD_CIRCULAR || w->model == CD_THREE_CIRCLE);
}

int check_ws(const cd_agents_t *a, const cd_workspace_t *w) {
  if (!ws_sizes_ok(w) || !w->ptr) return CD_ERR_ARG;
  if (w->n != a->n || w->model != a->model) return CD_ERR_WORKSPACE;
  if (w->bytes < make_layout(w).total) return CD_ERR_WORKSPACE;
  return CD_OK;
}

// threads per block of the force kernel (4 per agent)
inline int force_threads(const cd_workspace_t *w) {
  return w->force_threads > 0 ? (int)w->force_threads : 4 * kAgentsPerForceBlock;
}

template <typename T>
int launch_reduce(const Agents<T> &A, const Work &W, const Batch &B, int blocks,
                  bool clear_counts, hipStream_t s, int acc_slot = -1) {
  hipLaunchKernelGGL(k_reduce<T>, dim3(blocks), dim3(kBS), 0, s, A, W, B, clear_counts ? 1 : 0,
                     acc_slot);
  return check_launch();
}

inline void mark(hipEvent_t *ev, int k, hipStream_t s) {
  if (ev) (void)hipEventRecord(ev[k], s);
}

// Grid sizes of one step: per-scenario kernels (f) and force kernel (e).
struct Launch {
  int blocks_f, blocks_e, threads_e, blocks_t, lanes, scenarios;
};

Launch single_launch(const cd_workspace_t *w) {
  Launch L;
  L.blocks_f = single_parts(w->n);
  L.lanes = agent_lanes(w);
  L.blocks_e = std::max(1, grid_for(w->n, force_threads(w) / L.lanes));
  L.threads_e = force_threads(w);
  L.blocks_t = tile_blocks(w->cell_capacity, w);
  L.scenarios = 1;
  return L;
}

int check_step(const cd_agents_t *a, const cd_step_params_t *p, const cd_workspace_t *w) {
  int rc = check_agents(a);
  if (rc) return rc;
  if (!p) return CD_ERR_ARG;
  if ((p->flags & CD_USE_AGENT_AGENT) && !(p->cell_size > 0.0)) return CD_ERR_ARG;
  if (!(p->dt_min >= 0.0) || !(p->dt_max >= 0.0)) return CD_ERR_ARG;
  if ((p->flags & CD_USE_OBSTACLE) &&
      (p->n_segments < 0 || p->n_segments > 0x7fffffff || (p->n_segments > 0 && !p->segments)))
    return CD_ERR_ARG;
  if ((p->flags & CD_USE_NAV) &&
      (!p->nav_maps || p->nav_rows < 2 || p->nav_cols < 2 || p->nav_targets < 0 ||
       !(p->nav_step > 0.0) || (p->nav_mode != 0 && p->nav_mode != 1)))
    return CD_ERR_ARG;
  if (p->wall_cells && (p->wall_nx < 1 || p->wall_ny < 1 ||
                        p->wall_nx * p->wall_ny > ((int64_t)1 << 26) || !(p->wall_cell > 0.0)))
    return CD_ERR_ARG;
  return check_ws(a, w);
}

// ---------------------------------------------------------------------------
// Shared by the step kernels and the adjoint: the navigation lookup and the
// fused step's arguments.
// ---------------------------------------------------------------------------

// Navigation lookup in two halves: nav_fetch checks the cell and issues the
// map loads, nav_finish turns them into the direction (returns false when
// e0 is kept), so a kernel can put independent work between the two.
struct NavFetch {
  bool ok;
  int mode;
  double fx, fy, step;
  double p00, p01, p10, p11;  // potential (mode 1) or the direction in p00, p01 (mode 0)
};

__device__ __forceinline__ NavFetch nav_fetch(double px, double py, int64_t t, const double *maps,
                                              int64_t T_, int64_t rows, int64_t cols,
                                              double minx, double miny, double step, int mode) {
  NavFetch n;
  n.ok = false;
  n.mode = mode;
  n.step = step;
  n.fx = n.fy = n.p00 = n.p01 = n.p10 = n.p11 = 0.0;
  if (t < 0 || t >= T_) return n;
  const double qx = (px - minx) / step;
  const double qy = (py - miny) / step;
  if (mode == 0) {
    // indicer truncation + is_inside (quickest_path.py:41-44, navigation.py:59-78)
    if (!(qx > -1.0 && qx < (double)cols && qy > -1.0 && qy < (double)rows)) return n;
    const int64_t jj = (int64_t)qx, ii = (int64_t)qy;
    const double2 m = ((const double2 *)maps)[(t * rows + ii) * cols + jj];
    n.p00 = m.x;
    n.p01 = m.y;
    n.ok = true;
    return n;
  }
  if (!(qx >= 0.0 && qx <= (double)(cols - 1) && qy >= 0.0 && qy <= (double)(rows - 1))) return n;
  int64_t c0 = (int64_t)qx, r0 = (int64_t)qy;
  if (c0 > cols - 2) c0 = cols - 2;
  if (r0 > rows - 2) r0 = rows - 2;
  if (c0 < 0 || r0 < 0) return n;
  n.fx = qx - (double)c0;
  n.fy = qy - (double)r0;
  const double *P = maps + t * rows * cols;
  n.p00 = P[r0 * cols + c0];
  n.p01 = P[r0 * cols + c0 + 1];
  n.p10 = P[(r0 + 1) * cols + c0];
  n.p11 = P[(r0 + 1) * cols + c0 + 1];
  n.ok = true;
  return n;
}

__device__ __forceinline__ bool nav_finish(const NavFetch &n, double &ex, double &ey) {
  if (!n.ok) return false;
  if (n.mode == 0) {
    ex = n.p00;
    ey = n.p01;
    return true;
  }
  const double gx = ((n.p01 - n.p00) * (1.0 - n.fy) + (n.p11 - n.p10) * n.fy) / n.step;
  const double gy = ((n.p10 - n.p00) * (1.0 - n.fx) + (n.p11 - n.p01) * n.fx) / n.step;
  const double g = hypot_ref(gx, gy);
  if (!(g > 0.0)) return false;
  ex = -gx / g;
  ey = -gy / g;
  return true;
}

__device__ __forceinline__ bool nav_lookup(double px, double py, int64_t t, const double *maps,
                                           int64_t T_, int64_t rows, int64_t cols, double minx,
                                           double miny, double step, int mode, double &ex,
                                           double &ey) {
  return nav_finish(nav_fetch(px, py, t, maps, T_, rows, cols, minx, miny, step, mode), ex, ey);
}

struct FusedArgs {
  int flags, nav_mode;
  const double *maps;
  int64_t nav_T, nav_rows, nav_cols;
  double nav_minx, nav_miny, nav_step;
  const double4 *segs;
  double2 *tape_x;  // when set, the step also writes its new positions here ([n])
  void *tape_v;     // ... and its new velocities
  // wall grid (cd_wall_grid_build) or null: offsets of the nx * ny cells,
  // then the segment indices
  const int *wall_cells;
  int wall_nx, wall_ny;
  double wall_minx, wall_miny, wall_cell;
};

FusedArgs fused_args(const cd_step_params_t *p) {
  FusedArgs fa;
  fa.flags = p->flags;
  fa.nav_mode = p->nav_mode;
  fa.maps = p->nav_maps;
  fa.nav_T = p->nav_targets;
  fa.nav_rows = p->nav_rows;
  fa.nav_cols = p->nav_cols;
  fa.nav_minx = p->nav_minx;
  fa.nav_miny = p->nav_miny;
  fa.nav_step = p->nav_step;
  fa.segs = (const double4 *)p->segments;
  fa.tape_x = nullptr;
  fa.tape_v = nullptr;
  fa.wall_cells = p->wall_cells;
  fa.wall_nx = (int)p->wall_nx;
  fa.wall_ny = (int)p->wall_ny;
  fa.wall_minx = p->wall_minx;
  fa.wall_miny = p->wall_miny;
  fa.wall_cell = p->wall_cell;
  return fa;
}

}  // namespace
