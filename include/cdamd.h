/*
 * cdamd.h -- C ABI of the MI355X (gfx950) per-timestep agent update.
 *
 * All pointers in cd_agents_t and the workspace memory are DEVICE pointers
 * owned by the caller.  Every call is stream-ordered on `stream` (a
 * hipStream_t), performs no host synchronisation and no allocation, and may
 * be captured into a HIP graph, except the calls documented as synchronous
 * (cd_read_state, cd_batch_prepare, cd_batch_read_states, cd_step_timed,
 * cd_kernel_timed, cd_graph_*).  Calls return 0 on success or
 * a negative CD_ERR_* code; kernels record data-dependent conditions (grid
 * overflow, non-finite positions, full cell slot lists) as sticky bits in
 * the workspace state, read back by cd_read_state.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference repository root).
 */
#ifndef CDAMD_H
#define CDAMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *cd_stream_t; /* hipStream_t */

enum {
  CD_OK = 0,
  CD_ERR_ARG = -1,         /* null pointer, bad size or bad parameter */
  CD_ERR_TYPE = -2,        /* not a circular / three-circle agent array (InvalidType) */
  CD_ERR_WORKSPACE = -3,   /* workspace smaller than cd_workspace_bytes() or for other sizes */
  CD_ERR_UNSUPPORTED = -4, /* combination not built (e.g. the adjoint of three-circle agents) */
  CD_ERR_HIP = -5          /* a HIP runtime call or launch failed */
};

enum { CD_CIRCULAR = 0, CD_THREE_CIRCLE = 1 };
enum { CD_F64 = 0, CD_F32 = 1 };

/* Sticky status bits kept in the workspace state. */
enum {
  CD_STATUS_GRID_OVERFLOW = 1, /* grid larger than cell_capacity (or stale box): pairs skipped */
  CD_STATUS_NONFINITE = 2,     /* a non-finite position reached the binning */
  CD_STATUS_SLOTS_FULL = 4,    /* a cell held more than cell_slots agents: slow exact path used */
  CD_STATUS_NAV_KEPT = 8       /* adjoint or tangent step with navigation: an agent's lookup kept
                                  its previous e0 (outside the map), which was then taken from the
                                  agents' e0 and held constant */
};

/* Device structure-of-arrays view of agent_type_circular /
 * agent_type_three_circle records (crowddynamics/simulation/agents.py:299-307,
 * 447-448).  Vectors are interleaved [n][2].  `position`, `position_ls`,
 * `position_rs` and every scalar field are fp64; `velocity`,
 * `target_direction`, `force` and `force_prev` have the type named by `real`
 * (fp64, or fp32 in the fp32 mode).  Three-circle-only fields may be NULL for
 * circular agents. */
typedef struct cd_agents {
  int64_t n;
  int32_t model; /* CD_CIRCULAR / CD_THREE_CIRCLE */
  int32_t real;  /* CD_F64 / CD_F32 */
  double *position;
  void *velocity, *target_direction, *force, *force_prev;
  double *position_ls, *position_rs;
  const double *radius, *r_t, *r_s, *r_ts, *mass, *inertia_rot;
  const double *target_velocity, *target_angular_velocity;
  const double *tau_adj, *k_soc, *tau_0, *mu, *kappa, *damping, *std_rand_force;
  double *orientation, *angular_velocity, *target_orientation, *torque, *torque_prev;
  const double *tau_rot, *std_rand_torque;
  const int64_t *target;
} cd_agents_t;

/* Caller-owned device workspace and the sizes it was laid out for:
 * n agents, a binning grid of at most cell_capacity cells, cell_slots agent
 * slots per cell (cells holding more take an exact slow path). */
typedef struct cd_workspace {
  void *ptr;
  int64_t bytes;
  int64_t n;             /* agents (all scenarios of a batch together) */
  int64_t cell_capacity; /* grid cells (all scenarios of a batch together) */
  int64_t cell_slots;
  int32_t model;
  int32_t max_scenarios; /* scenario table size (0 or 1: single scenario) */
  /* set by cd_batch_prepare: */
  int32_t n_scenarios, blocks_f, blocks_e;
  int32_t force_threads; /* threads per block of the agent-parallel force kernel and the
                            adjoint: 64, 128 or 256 (0: 256) */
  int32_t tile_x, tile_y; /* force kernel: 0 = agent-parallel; else grid cells per
                             LDS tile, each 1..16 (same results either way) */
  int32_t blocks_t;       /* set by cd_batch_prepare */
  int32_t agent_lanes;    /* lanes per agent of the agent-parallel force kernel and the
                             adjoint: 2, 4, 8, 16, 32 or 64 (0: 4); same forces either way
                             (the adjoint's lane sums are in fp64) */
  int32_t order_every;    /* agent-parallel force kernel of the fused step, and the
                             adjoint kernel: 0 = agents in index order; k > 0 = in cell
                             order, rebuilt after the binning of every step whose counter
                             is a multiple of k (a wave's agents then share neighbour
                             cells); same results either way */
} cd_workspace_t;

/* One scenario of a batch: its agents are [agent_begin, agent_end) of the
 * batch arrays, its grid cells [cell_begin, cell_begin + cell_capacity) of
 * the workspace grid, its walls [seg_begin, seg_end) of params->segments.
 * Each scenario has its own adaptive dt and its own fluctuation key `seed`
 * (the noise counter is the agent's index within the scenario), so a batched
 * run reproduces the scenarios' standalone runs. */
typedef struct cd_scenario {
  int64_t agent_begin, agent_end;
  int64_t cell_begin, cell_capacity;
  int64_t seg_begin, seg_end;
  uint64_t seed;
  double dt_min, dt_max;
} cd_scenario_t;

/* Host copy of the workspace state scalars. */
typedef struct cd_state {
  double dt;          /* dt of the last integration (simulation.data['dt']) */
  double time_tot;    /* sum of dt (simulation.data['time_tot']) */
  double origin[2];   /* binning origin of the last binning */
  int64_t shape[2];   /* binning grid shape (nx, ny) of the last binning */
  int64_t status;     /* CD_STATUS_* bits */
  int64_t step;       /* steps taken by cd_step */
  double v_max;       /* max |v| of the last reduction */
  double v0_max;      /* max target_velocity of the last reduction */
} cd_state_t;

/* Offsets (bytes from the workspace base) of the block-list arrays that
 * cd_add_to_cells fills.  All int32. */
typedef struct cd_layout {
  int64_t points_indices; /* [n]: agents sorted stably by flat cell */
  int64_t cells_count;    /* [cell_capacity] */
  int64_t cells_offset;   /* [cell_capacity + 1]: exclusive prefix sum */
  int64_t cell_of;        /* [n]: flat cell of each agent */
  int64_t total_bytes;
} cd_layout_t;

/* ---- workspace --------------------------------------------------------- */

/* Bytes of workspace for the sizes in *ws (ptr and bytes are ignored).
 * Fills *layout when not NULL.  Negative for invalid sizes. */
int64_t cd_workspace_bytes(const cd_workspace_t *ws, cd_layout_t *layout);
/* Zeroes the workspace state and counters (once, before first use). */
int cd_workspace_init(const cd_workspace_t *ws, cd_stream_t stream);
/* Synchronous: copies the state scalars to *out. */
int cd_read_state(const cd_workspace_t *ws, cd_state_t *out, cd_stream_t stream);
/* Clears the sticky status bits (stream-ordered). */
int cd_clear_status(const cd_workspace_t *ws, cd_stream_t stream);

/* ---- block-list binning -------------------------------------------------- */

/* cell_lists.add_to_cells(points, cell_size) (called at
 * crowddynamics/core/interactions.py:192-193): counting sort of the agents
 * into square cells of side cell_size anchored at the positions' minimum.
 * Results in the workspace arrays named by cd_layout_t and the state's
 * origin/shape.  Deterministic: points_indices is the stable order. */
int cd_add_to_cells(const cd_agents_t *agents, double cell_size, const cd_workspace_t *ws,
                    cd_stream_t stream);

/* ---- LogicNode operators (crowddynamics/simulation/logic.py) ------------- */

/* agent_agent_block_list(agents, cell_size) (crowddynamics/core/interactions.py:191-205):
 * bins, then adds every adjacent pair's social + contact force (and torque)
 * to force (torque). */
int cd_agent_agent_block_list(const cd_agents_t *agents, double cell_size,
                              const cd_workspace_t *ws, cd_stream_t stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): cd_agent_agent_block_list with its workspace in a
 * library-owned device array of CD_NOALLOC_SCRATCH_BYTES, held on every GPU
 * the library runs on until the process ends.  The call allocates nothing.
 * cell_capacity and cell_slots size the grid as in cd_workspace_t; sizes
 * whose workspace (cd_workspace_bytes with max_scenarios 0) exceeds the
 * array return CD_ERR_WORKSPACE.  Run it with the stream's device current.
 * The call's CD_STATUS_* bits go to *status_out (a device int64, may be
 * NULL), written stream-ordered at the end of the call.  Every call on one
 * GPU shares its array: calls that overlap on one GPU (other streams, other
 * host threads) race and must be serialised by the caller. */
#define CD_NOALLOC_SCRATCH_BYTES (64ll << 20)
int cd_agent_agent_block_list_noalloc(const cd_agents_t *agents, double cell_size,
                                      int64_t cell_capacity, int64_t cell_slots,
                                      int64_t *status_out, cd_stream_t stream);

/* agent_obstacle(agents, obstacles) (crowddynamics/core/interactions.py:208-214):
 * segments is obstacle_type_linear[W] = {p0: f8[2], p1: f8[2]}, [W][4] fp64. */
int cd_agent_obstacle(const cd_agents_t *agents, const double *segments, int64_t n_segments,
                      cd_stream_t stream);

/* force_adjust_agents(agents) (crowddynamics/core/motion/adjusting.py:98-108) */
int cd_force_adjust_agents(const cd_agents_t *agents, cd_stream_t stream);
/* torque_adjust_agents(agents) (crowddynamics/core/motion/adjusting.py:111-121) */
int cd_torque_adjust_agents(const cd_agents_t *agents, cd_stream_t stream);
/* orient_towards_target_direction(agents) (crowddynamics/core/steering/orientation.py:17-21) */
int cd_orient_towards_target_direction(const cd_agents_t *agents, cd_stream_t stream);

/* Navigation.update (crowddynamics/simulation/logic.py:149-165) with
 * meshgrid().indicer (core/steering/quickest_path.py:41-44) and getdefault
 * (core/steering/navigation.py:67-78).  mode 0: `maps` = n_targets direction
 * maps [rows][cols][2] read at the truncated index (reference).  mode 1:
 * `maps` = n_targets potentials [rows][cols], e0 = -grad/|grad| of the
 * bilinear interpolant.  Agents whose target is outside [0, n_targets) keep
 * their direction. */
int cd_navigation(const cd_agents_t *agents, const double *maps, int64_t n_targets,
                  int64_t rows, int64_t cols, double minx, double miny, double step,
                  int32_t mode, cd_stream_t stream);

/* Fluctuation.update (crowddynamics/simulation/logic.py:78-86) with
 * force_fluctuation / torque_fluctuation (core/motion/fluctuation.py:14-62):
 * adds the step's noise drawn from a Philox4x32-10 stream keyed by seed and
 * counted by (agent, step). */
int cd_force_fluctuation(const cd_agents_t *agents, uint64_t seed, uint64_t step,
                         cd_stream_t stream);

/* velocity_verlet_integrator(agents, dt_min, dt_max)
 * (crowddynamics/core/integrator.py:209-256): adaptive dt (:29-97) on the
 * device, then the reference's translational (and rotational + shoulders)
 * update.  dt and time_tot are kept in the workspace state. */
int cd_velocity_verlet_integrator(const cd_agents_t *agents, double dt_min, double dt_max,
                                  const cd_workspace_t *ws, cd_stream_t stream);

/* Reset.update (crowddynamics/simulation/logic.py:59-64) */
int cd_reset(const cd_agents_t *agents, cd_stream_t stream);

/* ---- fused step ---------------------------------------------------------- */

enum {
  CD_USE_FLUCT = 1,
  CD_USE_NAV = 2,
  CD_USE_ADJUST = 4,        /* Adjusting (+ Orientation for three-circle) */
  CD_USE_AGENT_AGENT = 8,
  CD_USE_OBSTACLE = 16,
  CD_USE_RESET = 32,
  /* Orientation: target_orientation = angle(velocity) instead of the
   * reference's angle(target_direction) (core/steering/orientation.py:17-21);
   * three-circle agents, fused step only */
  CD_USE_ORIENT_VELOCITY = 64
};

typedef struct cd_step_params {
  double cell_size, dt_min, dt_max;
  int32_t flags;    /* CD_USE_* */
  int32_t nav_mode; /* as cd_navigation */
  uint64_t seed;
  const double *segments;
  int64_t n_segments;
  const double *nav_maps;
  int64_t nav_targets, nav_rows, nav_cols;
  double nav_minx, nav_miny, nav_step;
  /* Optional wall grid of the fused step of a single scenario (device
   * memory written by cd_wall_grid_build and copied to the device): the
   * segments binned into square cells, so that an agent visits only the
   * segments within its reach.  NULL: every agent scans every segment.
   * Same results either way. */
  const int32_t *wall_cells;
  int64_t wall_nx, wall_ny;
  double wall_minx, wall_miny, wall_cell;
} cd_step_params_t;

/* Host only (no device call): bins segments ([W][4] fp64, HOST memory) into
 * square cells of side `cell` over their bounding box grown by `reach`, the
 * largest distance from an agent's centre at which a wall contact can occur
 * (circular: the radius; three-circle: r_ts + max(r_t, r_s)).  A cell lists
 * every segment whose bounding box lies within reach of it, so an agent in
 * the cell is within reach of no other segment, and an agent outside the grid
 * of none.  Writes into out[0 .. capacity) the offsets of the nx * ny cells
 * ([nx * ny + 1], flat index ix * ny + iy) followed by the segment indices,
 * and into *params the grid fields (wall_nx, wall_ny, wall_minx, wall_miny,
 * wall_cell; the caller sets wall_cells to the device copy of out).  Returns
 * the number of int32 the grid needs (call with out NULL to size it), or a
 * negative CD_ERR_* code.  Replaces the per-step O(N W) scan of
 * agent_obstacle (crowddynamics/core/interactions.py:169-186) inside the
 * fused step. */
int64_t cd_wall_grid_build(const double *segments, int64_t n_segments, double reach,
                           double cell, int32_t *out, int64_t capacity,
                           cd_step_params_t *params);

/* One MultiAgentSimulation.update() (crowddynamics/simulation/multiagent.py:51-55)
 * of the tree Reset << (Integrator << (Fluctuation, Adjusting << (Navigation,
 * Orientation), AgentAgentInteractions, AgentObstacleInteractions))
 * (crowddynamics/examples/simulations.py:123-136) in two kernels: the
 * per-agent terms (fluctuation, navigation, orientation, adjusting) with the
 * binning, then the pair and wall forces with the integration.  The
 * fluctuation step counter and every reduction stay on the device.
 * cd_step_prepare must run once before the first cd_step (and after any
 * host-side change of positions or velocities). */
int cd_step_prepare(const cd_agents_t *agents, const cd_step_params_t *params,
                    const cd_workspace_t *ws, uint64_t first_step, cd_stream_t stream);
int cd_step(const cd_agents_t *agents, const cd_step_params_t *params, const cd_workspace_t *ws,
            cd_stream_t stream);
/* cd_step that also writes the new positions to tape_x ([n][2] fp64) and,
 * when tape_v is not null, the new velocities to tape_v ([n][2] in the
 * agents' real type): a trajectory recorded by the step itself, with no
 * copies.  No reference counterpart (the rollout drivers below use it). */
int cd_step_tape(const cd_agents_t *agents, const cd_step_params_t *params,
                 const cd_workspace_t *ws, double *tape_x, void *tape_v, cd_stream_t stream);

/* Batched ensemble: independent scenarios advanced together, one launch per
 * step kernel over all scenarios (two launches per step, graph-capturable).
 * Synchronous: uploads the scenario table and block maps into the
 * workspace (which needs max_scenarios >= n_scenarios) and fills ws->
 * n_scenarios / blocks_f / blocks_e.  params supplies the cell size, flags,
 * the concatenated wall segments and the navigation maps shared by all
 * scenarios; its seed, dt_min and dt_max are replaced per scenario.  No
 * reference counterpart (the reference runs one simulation per process). */
int cd_batch_prepare(const cd_agents_t *agents, const cd_step_params_t *params,
                     const cd_scenario_t *scenarios, int32_t n_scenarios, cd_workspace_t *ws,
                     uint64_t first_step, cd_stream_t stream);
int cd_batch_step(const cd_agents_t *agents, const cd_step_params_t *params,
                  const cd_workspace_t *ws, cd_stream_t stream);
/* Synchronous: state scalars of scenarios [0, count). */
int cd_batch_read_states(const cd_workspace_t *ws, cd_state_t *out, int32_t count,
                         cd_stream_t stream);

/* Adjoint of one fused step (circular agents, fp64, every node of the
 * standard tree with RESET set: fluctuation, navigation, adjusting, agent
 * and wall interactions, fixed or adaptive dt; single-scenario workspace).
 * `agents` holds the state x_k, v_k the step started from (position /
 * velocity may point into a caller's tape).  Maps the adjoint (lam_x, lam_v,
 * lam_g) of step k+1 to that of step k in place, adds grad_x (dL/dx_k,
 * [n][2], may be null) to lam_x, and adds the step's parameter gradients to
 * grad ([n][5]: tau_0, mu, kappa, damping, target_velocity of each agent).
 * With navigation, e0 is looked up again at x_k (its derivative in the
 * bilinear mode; zero at the nearest node); an agent whose lookup keeps its
 * e0 sets CD_STATUS_NAV_KEPT.  With adaptive dt (dt_min < dt_max), adj->
 * x_next / v_next give state k + 1, and dt's derivative reaches the agent of
 * the largest speed and the one of the largest target_velocity while dt lies
 * between its clamps.  Stream-ordered, no allocation.  Replaces
 * reverse-mode differentiation through interactions.py:51-70 and 107-141,
 * adjusting.py:16-51, navigation.py:59-78 and integrator.py:29-97 and
 * 164-176, which the reference does not provide. */
typedef struct cd_adjoint {
  double *lam_x, *lam_v, *lam_g; /* [n][2] adjoint state, in/out */
  double *lam_f, *lam_vt;        /* [n][2] scratch of the step */
  double *grad;                  /* [n][5] accumulated parameter gradients */
  /* [n][2] position and velocity of step k + 1: needed with adaptive dt
   * (dt_min < dt_max), whose derivative enters through them; else unused */
  const double *x_next, *v_next;
} cd_adjoint_t;
int cd_adjoint_step(const cd_agents_t *agents, const cd_step_params_t *params,
                    const double *grad_x, const cd_adjoint_t *adj, const cd_workspace_t *ws,
                    cd_stream_t stream);

/* Tangent (forward-mode) step, the scope of cd_adjoint_step: from the
 * state x_k, v_k in `agents`, maps the tangent (dx, dv, dg) of (position,
 * velocity, force_prev) at step k to step k+1 in place, along the parameter
 * tangent dp ([n][5]: tau_0, mu, kappa, damping, target_velocity); df
 * ([n][2]) receives the tangent of the step's force.  With cd_adjoint_step
 * it gives the dot-product test <J t, u> = <t, J^T u> of a rollout.
 * Stream-ordered, no allocation.  No reference counterpart (the reference
 * has no derivatives of interactions.py:51-70, adjusting.py:16-51 or
 * integrator.py:164-176). */
typedef struct cd_tangent {
  double *dx, *dv, *dg; /* [n][2] tangent state, in/out */
  double *df;           /* [n][2] out: tangent of the step's force */
  const double *dp;     /* [n][5] parameter tangent */
  const double *x_next, *v_next; /* [n][2] state k + 1, as in cd_adjoint_t */
} cd_tangent_t;
int cd_tangent_step(const cd_agents_t *agents, const cd_step_params_t *params,
                    const cd_tangent_t *tan, const cd_workspace_t *ws, cd_stream_t stream);

/* Differentiable rollout drivers (circular, fp64, the cd_adjoint_step scope).
 * cd_rollout_record: from the agents' current state (after cd_step_prepare),
 * `steps` fused steps; positions to x_out [steps + 1][n][2] (x_out[0] the
 * start) and, before steps k = 0, checkpoint, 2 * checkpoint, ..., the
 * state (position, velocity, force_prev) to ck_* [k / checkpoint][n][2].
 * cd_adjoint_segment: restores checkpoint state ck_*[k0] (the caller passes
 * the pointers of that checkpoint), re-runs steps k0 .. k1 - 2 into
 * tape_x / tape_v [k1 - k0][n][2] (bit-identical to the recorded run; with
 * adaptive dt also step k1 - 1, so the tapes hold k1 - k0 + 1 states), then
 * applies cd_adjoint_step for k = k1 - 1 down to k0 with grad_x + k * n * 2
 * (dL/dx_k of the whole rollout, may be null).  Stream-ordered; no reference
 * counterpart (reverse-mode differentiation of the update). */
int cd_rollout_record(const cd_agents_t *agents, const cd_step_params_t *params,
                      const cd_workspace_t *ws, int64_t steps, int64_t checkpoint,
                      double *x_out, double *ck_pos, double *ck_vel, double *ck_fprev,
                      cd_stream_t stream);
int cd_adjoint_segment(const cd_agents_t *agents, const cd_step_params_t *params,
                       const cd_workspace_t *ws, int64_t k0, int64_t k1, const double *ck_pos,
                       const double *ck_vel, const double *ck_fprev, double *tape_x,
                       double *tape_v, const double *grad_x, const cd_adjoint_t *adj,
                       cd_stream_t stream);

/* Synchronous: runs `steps` steps (cd_step, or cd_batch_step on a workspace
 * prepared by cd_batch_prepare) and returns in kernel_ms[4] the summed GPU
 * time of the first kernel (per-agent terms and binning), the force kernel
 * (forces and integration), 0 (no separate integration kernel) and the
 * whole step, measured with HIP events recorded on `stream` around each
 * launch.  No reference counterpart (measurement). */
int cd_step_timed(const cd_agents_t *agents, const cd_step_params_t *params,
                  const cd_workspace_t *ws, int64_t steps, double *kernel_ms,
                  cd_stream_t stream);

/* Synchronous measurement of one step kernel (0: the first kernel, per-agent
 * terms and binning; 1: the force kernel, forces and integration): after one
 * pass of the first kernel, `reps` back-to-back launches of it
 * between two HIP events on `stream`; *avg_ms = elapsed / reps.  Repeating
 * a kernel leaves the agent state and workspace meaningless (forces
 * accumulate, agents move, cell counts grow): use a copy, then discard it
 * or prepare again.  No reference counterpart (measurement). */
int cd_kernel_timed(const cd_agents_t *agents, const cd_step_params_t *params,
                    const cd_workspace_t *ws, int32_t kernel, int64_t reps, double *avg_ms,
                    cd_stream_t stream);

/* ---- navigation field construction --------------------------------------- */

/* Bytes of device scratch cd_distance_map needs for a rows x cols node grid
 * (negative for invalid sizes). */
int64_t cd_field_work_bytes(int64_t rows, int64_t cols);

/* distance_map (crowddynamics/core/steering/quickest_path.py:54-115), i.e.
 * scikit-fmm 0.0.9's distance(phi, dx) on the node grid: phi [rows][cols]
 * fp64 (row = y, col = x; > 0 target region, < 0 elsewhere), mask
 * [rows][cols] uint8 (non-zero: excluded node, may be NULL), spacing dx.
 * out [rows][cols]: the signed distance to the zero contour (negative where
 * phi < 0), NaN at excluded and unreachable nodes; nodes next to the
 * contour by scikit-fmm's linear interpolation, the rest by the first-order
 * upwind Eikonal update.  Synchronous (the relaxation passes read a
 * convergence flag); work is device scratch of cd_field_work_bytes bytes.
 * CD_ERR_HIP also when the passes do not converge within their bound. */
int cd_distance_map(const double *phi, const uint8_t *mask, int64_t rows, int64_t cols,
                    double dx, double *out, void *work, int64_t work_bytes,
                    cd_stream_t stream);

/* direction_map (quickest_path.py:125-149): the normalised np.gradient of
 * dmap [rows][cols], as ex (x, along columns) and ey (y, along rows); NaN
 * where the gradient vanishes or a NaN enters it. */
int cd_direction_map(const double *dmap, int64_t rows, int64_t cols, double *ex, double *ey,
                     cd_stream_t stream);

/* obstacle_handling (core/steering/obstacle_handling.py:15-74) over n nodes:
 * where 0 < -dobs < radius, (tx, ty) <- -p (ox, oy) + (1 - p) (tx, ty) with
 * p = strength^(-dobs / radius); then every node is normalised.  dobs and
 * (ox, oy) are the obstacles' distance and direction maps. */
int cd_obstacle_handling(const double *dobs, const double *ox, const double *oy, double *tx,
                         double *ty, int64_t n, double radius, double strength,
                         cd_stream_t stream);

/* fill_missing (quickest_path.py:152-166): (ex, ey) = (sx, sy), except that
 * nodes with fill[k] != 0 whose direction is NaN take the direction of the
 * nearest node (index distance, within `reach` nodes; ties to the smaller
 * row, then column) whose direction is not NaN.  ex, ey must not alias sx,
 * sy. */
int cd_fill_missing(const double *sx, const double *sy, const uint8_t *fill, int64_t rows,
                    int64_t cols, int32_t reach, double *ex, double *ey, cd_stream_t stream);

/* ---- agent state nodes and herding ---------------------------------------- */

/* InsideDomain.update (crowddynamics/simulation/logic.py:351-357): active[i]
 * (uint8, the agents' `active` flags) = position i inside the polygon
 * vertices [V][2] (fp64, device; closed from the last vertex back to the
 * first), by the crossings test of matplotlib's Path.contains_points, which
 * the reference calls; *inactive (a device uint64) += the number of agents
 * whose flag changed (data['inactive']). */
int cd_inside_domain(const double *position, int64_t n, const double *vertices,
                     int64_t n_vertices, uint8_t *active, uint64_t *inactive,
                     cd_stream_t stream);
/* TargetReached.update (logic.py:383-387) for one polygon target:
 * reached_by[i] |= position i inside the polygon; *count (device uint64) =
 * the number of agents with reached_by set (data['target_<i>']). */
int cd_target_reached(const double *position, int64_t n, const double *vertices,
                      int64_t n_vertices, uint8_t *reached_by, uint64_t *count,
                      cd_stream_t stream);
/* find_nearest_neighbors (core/steering/collective_motion.py:69-107, with
 * the block list at cell size = sight, :250-257): neighbors [n][k] (int64,
 * device) = the k (1..16) nearest agents strictly within sight whose line of
 * sight crosses none of the segments ([W][4], core/geom2D.py:38-59
 * line_intersect), nearest first (ties to the smaller index), -1 where
 * missing.  Bins the agents into the workspace (cd_add_to_cells). */
int cd_nearest_neighbors(const cd_agents_t *agents, double sight, int32_t k,
                         const double *segments, int64_t n_segments, int64_t *neighbors,
                         const cd_workspace_t *ws, cd_stream_t stream);
/* herding_interaction (collective_motion.py:110-154): for agents with
 * is_follower[i] set, the normalised weighted average of the direction to
 * the mean position and of the mean velocity of the neighbours heading away
 * (herding_relationship, :26-58, with cos_phi = cos(phi)); direction [n][2]
 * (zero otherwise) and has_direction [n] (uint8). */
int cd_herding_interaction(const double *position, const double *velocity, int64_t n,
                           const uint8_t *is_follower, const int64_t *neighbors, int32_t k,
                           double weight_position, double cos_phi, double *direction,
                           uint8_t *has_direction, cd_stream_t stream);

/* ---- HIP graphs (synchronous helpers) ------------------------------------- */

int cd_graph_begin(cd_stream_t stream);
int cd_graph_end(cd_stream_t stream, void **graph_exec);
int cd_graph_launch(void *graph_exec, cd_stream_t stream);
int cd_graph_destroy(void *graph_exec);

/* ---- misc ----------------------------------------------------------------- */

/* Philox4x32-10 on the device for one counter (testing the noise stream):
 * in = {ctr0..3, key0, key1}, out = 4 words. */
int cd_philox4x32(const uint32_t *in_dev, uint32_t *out_dev, cd_stream_t stream);
/* The device's length of (x, y) for n pairs xy ([n][2] fp64, device) into
 * out ([n], device): the sequence of the C library's hypot that np.hypot
 * reaches (core/vector2D.py:113), bit for bit (testing the lengths every
 * distance, normal and max |v| of the step takes). */
int cd_hypot(const double *xy, double *out, int64_t n, cd_stream_t stream);
const char *cd_version(void);

#ifdef __cplusplus
}
#endif

#endif /* CDAMD_H */
