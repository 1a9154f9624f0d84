// Adjoint of the fused step (circular agents, fp64, fixed dt): the reverse
// sweep that differentiates a trajectory loss with respect to the agents'
// tau_0, mu, kappa, damping and target_velocity.
//
// Forward step k in the reference's order (core/integrator.py:164-176,
// crowddynamics/simulation/logic.py:59-75; G is force_prev):
//   F_k     = adjust(v_k; v0) + sum_j f(x_i - x_j, v_i - v_j; p_i) + noise_k
//   v_{k+1} = v_k + (G_k + F_k) / (2 m) * dt
//   x_{k+1} = x_k + v_{k+1} * dt + F_k / (2 m) * dt^2
//   G_{k+1} = F_k
// Given the adjoints (lx, lv, lg) of step k+1 the sweep forms
//   lvt = lv + dt lx,   lF = lg + dt/(2m) lvt + dt^2/(2m) lx,
//   lg_k = dt/(2m) lvt,   lv_k = lvt + (dF/dv)^T lF,   lx_k = lx + (dF/dx)^T lF,
//   grad_p += (dF/dp)^T lF,
// and the caller adds dL/dx_k.  The noise does not depend on the state or
// the parameters; Reset (logic.py:59-64) makes F_k the only force term.
//
// Pair terms are owner-computes like the forward force kernel: agent i
// collects, for every neighbour j, J(d; p_i)^T lF_i from its own side and
// -J(d; p_j)^T lF_j from j's side (the pair function is odd in
// d = (x_i - x_j, v_i - v_j), so j's side is -f(d; p_j) and its Jacobian
// with respect to x_i is -J(d; p_j)).  The 2x4 Jacobian J and the
// parameter derivatives come from evaluating the forward pair function on
// dual numbers, which keeps every branch and cut-off of the step (sight,
// tau validity, force truncation, contact).  No atomics: each agent writes
// only its own adjoint and gradient rows.

#include "cd_common.h"
#include "cd_dual.h"

namespace {

struct AdjView {
  double2 *lam_x, *lam_v, *lam_g;  // adjoint state, [n]
  double2 *lam_f, *lam_vt;         // lF and lvt of the current step, [n]
  double *grad;                    // [n][5]: tau_0, mu, kappa, damping, target_velocity
  const double2 *grad_x;           // dL/dx_k or null
};

__global__ void __launch_bounds__(256) k_adj_pre(int n, const double *mass, double dt,
                                                 AdjView J) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double c = dt / (2.0 * mass[i]);
  const double2 lx = J.lam_x[i], lv = J.lam_v[i], lg = J.lam_g[i];
  const double2 lvt = make_double2(lv.x + dt * lx.x, lv.y + dt * lx.y);
  const double c2 = dt * c;
  J.lam_f[i] = make_double2(lg.x + c * lvt.x + c2 * lx.x, lg.y + c * lvt.y + c2 * lx.y);
  J.lam_vt[i] = lvt;
  J.lam_g[i] = make_double2(c * lvt.x, c * lvt.y);
}

using D4 = Dual<4>;

// Accumulates into acc[0..7] (lx.x, lx.y, lv.x, lv.y, tau_0, mu, kappa,
// damping) the contributions of agent i's pair with j, d = (x, y, vx, vy).
__device__ __forceinline__ void pair_vjp(double x, double y, double vx, double vy, double r_tot,
                                         const Owner<double> &oi, const Owner<double> &oj,
                                         double2 lfi, double2 lfj, double acc[8]) {
  {  // i's side: derivatives along the relative state
    const Owner<D4> o{D4(oi.neg_mk), D4(oi.tau_0), D4(oi.mu), D4(oi.kappa), D4(oi.damping)};
    D4 fx, fy;
    if (!pair_circular<D4>(D4::seed(x, 0), D4::seed(y, 1), D4::seed(vx, 2), D4::seed(vy, 3),
                           D4(r_tot), o, fx, fy))
      return;  // beyond sight: neither side contributes
#pragma unroll
    for (int k = 0; k < 4; ++k) acc[k] += lfi.x * fx.d[k] + lfi.y * fy.d[k];
  }
  {  // i's side: derivatives along i's parameters
    const Owner<D4> o{D4(oi.neg_mk), D4::seed(oi.tau_0, 0), D4::seed(oi.mu, 1),
                      D4::seed(oi.kappa, 2), D4::seed(oi.damping, 3)};
    D4 fx, fy;
    pair_circular<D4>(D4(x), D4(y), D4(vx), D4(vy), D4(r_tot), o, fx, fy);
#pragma unroll
    for (int k = 0; k < 4; ++k) acc[4 + k] += lfi.x * fx.d[k] + lfi.y * fy.d[k];
  }
  {  // j's side, -f(d; p_j)
    const Owner<D4> o{D4(oj.neg_mk), D4(oj.tau_0), D4(oj.mu), D4(oj.kappa), D4(oj.damping)};
    D4 fx, fy;
    pair_circular<D4>(D4::seed(x, 0), D4::seed(y, 1), D4::seed(vx, 2), D4::seed(vy, 3),
                      D4(r_tot), o, fx, fy);
#pragma unroll
    for (int k = 0; k < 4; ++k) acc[k] -= lfj.x * fx.d[k] + lfj.y * fy.d[k];
  }
}

__device__ __forceinline__ Owner<double> owner_at(const Agents<double> &A, int j) {
  Owner<double> o;
  o.neg_mk = (-A.mass[j]) * A.k_soc[j];
  o.tau_0 = A.tau_0[j];
  o.mu = A.mu[j];
  o.kappa = A.kappa[j];
  o.damping = A.damping[j];
  return o;
}

// Calls fn(j, x_j, v_j, r_j) for part q (of L equal parts) of the
// candidates of agent i's 3 x 3 stencil, taken as one sequence (rows
// dx = -1, 0, +1, each row's cells dy = -1, 0, +1, each cell's slots in
// agent order), skipping i itself.
template <class Fn>
__device__ __forceinline__ void stencil_each(const Agents<double> &A, const Work &W, int c0,
                                             int nx, int ny, int lx, int ly, int q, int L,
                                             int a0, int a1, int i, Fn &&fn) {
  int gr[3], n[9];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const int ox = lx + r - 1;
    const bool okx = ox >= 0 && ox < nx;
    gr[r] = c0 + ox * ny + ly - 1;
    n[3 * r] = (okx && ly - 1 >= 0) ? W.count[gr[r]] : 0;
    n[3 * r + 1] = okx ? W.count[gr[r] + 1] : 0;
    n[3 * r + 2] = (okx && ly + 1 < ny) ? W.count[gr[r] + 2] : 0;
  }
  int pre[10];
  pre[0] = 0;
#pragma unroll
  for (int k = 0; k < 9; ++k) pre[k + 1] = pre[k] + n[k];
  const int tot = pre[9];
  // floor(tot q / L) by a shift (L is a power of two), in 64 bits: the
  // counts of overfull cells are not bounded by the slots
  const int lg = __builtin_ctz(L);
  const int b0 = (int)(((long long)tot * q) >> lg);
  const int b1 = (int)(((long long)tot * (q + 1)) >> lg);
  const double2 *e_vel = (const double2 *)W.e_vel;
  bool over = false;
#pragma unroll
  for (int k = 0; k < 9; ++k) over = over || n[k] > W.K;
  if (!over) {  // one copy of the pair code: the candidates in order
    for (int idx = b0; idx < b1; ++idx) {
      int base = 0, c = gr[0];
#pragma unroll
      for (int k = 1; k < 9; ++k) {
        if (idx >= pre[k]) {
          base = pre[k];
          c = gr[k / 3] + k % 3;
        }
      }
      const size_t e = (size_t)c * W.K + (idx - base);
      const int j = W.ell_ids[e];
      if (j == i) continue;
      fn(j, W.e_pos[e], e_vel[e], W.e_r[e]);
    }
    return;
  }
  // a cell with more agents than slots (exact slow path): members in agent
  // order, cell by cell
#pragma unroll 1
  for (int k = 0; k < 9; ++k) {
    const int cnt = n[k];
    const int s0 = max(0, b0 - pre[k]), s1 = min(cnt, b1 - pre[k]);
    if (s0 >= s1) continue;
    const int c = gr[k / 3] + k % 3;
    if (cnt <= W.K) {  // an ordinary cell: its slots
      for (int s2 = s0; s2 < s1; ++s2) {
        const size_t e = (size_t)c * W.K + s2;
        const int j = W.ell_ids[e];
        if (j == i) continue;
        fn(j, W.e_pos[e], e_vel[e], W.e_r[e]);
      }
      continue;
    }
    int r = 0;
    for (int j = a0; j < a1 && r < s1; ++j) {
      if (W.cell_of[j] != c) continue;
      if (r >= s0 && j != i) {
        const V2<double> vj = A.vel[j];
        fn(j, A.pos[j], make_double2(vj.x, vj.y), A.radius[j]);
      }
      ++r;
    }
  }
}

// Pair adjoints of agent i with part q (of L equal parts) of its stencil's
// candidates.
__device__ __forceinline__ void stencil_vjp(const Agents<double> &A, const Work &W,
                                            const AdjView &J, int c0, int nx, int ny, int lx,
                                            int ly, int q, int L, int a0, int a1, int i,
                                            double2 xi, double2 vi, double ri,
                                            const Owner<double> &oi, double2 lfi,
                                            double acc[8]) {
  stencil_each(A, W, c0, nx, ny, lx, ly, q, L, a0, a1, i,
               [&](int j, double2 xj, double2 vj, double rj) {
                 pair_vjp(xi.x - xj.x, xi.y - xj.y, vi.x - vj.x, vi.y - vj.y, ri + rj, oi,
                          owner_at(A, j), lfi, J.lam_f[j], acc);
               });
}

// L lanes per agent (2 to 64) as in the forward force kernel: lane q
// takes part q of the stencil's candidates; lane 0 adds the agent's own
// (adjusting) terms and writes the results.
__global__ void __launch_bounds__(256) k_adjoint(Agents<double> A, Work W, Batch B, AdjView J,
                                                 int flags, int L) {
  const int2 sk = block_e(B);
  const ScnDesc d = desc_of(B, sk.x);
  const ScnState *ss = &B.sst[sk.x];
  const int lg = __builtin_ctz(L);  // L is a power of two
  // the agent: the next in index order, or in the forward steps' processing
  // order (any permutation gives the same sums: each agent's lanes combine
  // in a fixed order and agents share no accumulator)
  int i = d.a0 + sk.y * ((int)blockDim.x >> lg) + ((int)threadIdx.x >> lg);
  if (W.order_every > 0 && i < d.a1) i = W.order[i];
  const int q = threadIdx.x & (L - 1);
  const bool have = i >= d.a0 && i < d.a1;
  const bool binned = (flags & CD_USE_AGENT_AGENT) && ss->valid && !ss->oob[0];
  double acc[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) acc[k] = 0.0;
  double2 lfi = make_double2(0.0, 0.0);
  if (have) lfi = J.lam_f[i];
  if (have && binned) {
    const double2 xi = A.pos[i];
    const V2<double> v = A.vel[i];
    const double2 vi = make_double2(v.x, v.y);
    const int c = W.cell_of[i] - d.c0;
    const int ny = ss->ny, nx = ss->nx;
    const int lx = c / ny, ly = c - lx * ny;
    stencil_vjp(A, W, J, d.c0, nx, ny, lx, ly, q, L, d.a0, d.a1, i, xi, vi, A.radius[i],
                owner_at(A, i), lfi, acc);
  }
  // sums over the L lanes (butterfly in a fixed order: deterministic)
  for (int mask = 1; mask < L; mask <<= 1) {
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] += __shfl_xor(acc[k], mask, 64);
  }
  double sum[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) sum[k] = acc[k];
  if (!have || q != 0) return;
  double2 lv = J.lam_vt[i];
  double g_v0 = 0.0;
  if (flags & CD_USE_ADJUST) {
    // force_adjust (core/motion/adjusting.py:16-51): F = (m / tau_adj) (v0 e0 - v)
    const double s = A.mass[i] / A.tau_adj[i];
    const V2<double> e = A.e0[i];
    lv.x -= s * lfi.x;
    lv.y -= s * lfi.y;
    g_v0 = s * (e.x * lfi.x + e.y * lfi.y);
  }
  double2 lx = J.lam_x[i];
  lx.x += sum[0];
  lx.y += sum[1];
  if (J.grad_x) {
    const double2 g = J.grad_x[i];
    lx.x += g.x;
    lx.y += g.y;
  }
  lv.x += sum[2];
  lv.y += sum[3];
  J.lam_x[i] = lx;
  J.lam_v[i] = lv;
  double *g = J.grad + (size_t)i * 5;
  g[0] += sum[4];
  g[1] += sum[5];
  g[2] += sum[6];
  g[3] += sum[7];
  g[4] += g_v0;
}

// ---------------------------------------------------------------------------
// Tangent (forward-mode) step: the directional derivative of the same step
// along a tangent of the state (dx, dv, dg = d force_prev) and of the
// parameters (dp).  With the adjoint it forms the dot-product test
// <J t, u> = <t, J^T u> of the whole rollout map, which holds at any
// horizon for exact derivatives of the same floating-point program.
//   dF_k     = (dF/dx) dx_k + (dF/dv) dv_k + (dF/dp) dp
//   dv_{k+1} = dv_k + (dG_k + dF_k) dt / (2 m)
//   dx_{k+1} = dx_k + dv_{k+1} dt + dF_k dt^2 / (2 m)
//   dG_{k+1} = dF_k
// Agent i's pair terms are f(x_i - x_j, v_i - v_j; p_i), evaluated on
// one-direction dual numbers seeded with (dx_i - dx_j, dv_i - dv_j) and
// i's parameter tangents, so every branch is the forward step's.
// ---------------------------------------------------------------------------

struct TanView {
  double2 *dx, *dv, *dg;  // tangent state, [n]
  double2 *df;            // dF_k of the current step, [n]
  const double *dp;       // [n][5]: tau_0, mu, kappa, damping, target_velocity
};

using D1 = Dual<1>;

__device__ __forceinline__ D1 d1(double v, double t) {
  D1 r(v);
  r.d[0] = t;
  return r;
}

__global__ void __launch_bounds__(256) k_tangent(Agents<double> A, Work W, Batch B, TanView T,
                                                 int flags, int L) {
  const int2 sk = block_e(B);
  const ScnDesc d = desc_of(B, sk.x);
  const ScnState *ss = &B.sst[sk.x];
  const int lg = __builtin_ctz(L);
  int i = d.a0 + sk.y * ((int)blockDim.x >> lg) + ((int)threadIdx.x >> lg);
  if (W.order_every > 0 && i < d.a1) i = W.order[i];
  const int q = threadIdx.x & (L - 1);
  const bool have = i >= d.a0 && i < d.a1;
  const bool binned = (flags & CD_USE_AGENT_AGENT) && ss->valid && !ss->oob[0];
  double acc[2] = {0.0, 0.0};
  if (have && binned) {
    const double2 xi = A.pos[i];
    const V2<double> v = A.vel[i];
    const double2 vi = make_double2(v.x, v.y);
    const double2 dxi = T.dx[i], dvi = T.dv[i];
    const double ri = A.radius[i];
    const double *dpi = T.dp + (size_t)i * 5;
    const Owner<D1> oi{D1((-A.mass[i]) * A.k_soc[i]), d1(A.tau_0[i], dpi[0]),
                       d1(A.mu[i], dpi[1]), d1(A.kappa[i], dpi[2]), d1(A.damping[i], dpi[3])};
    const int c = W.cell_of[i] - d.c0;
    const int ny = ss->ny, nx = ss->nx;
    const int lx = c / ny, ly = c - lx * ny;
    stencil_each(A, W, d.c0, nx, ny, lx, ly, q, L, d.a0, d.a1, i,
                 [&](int j, double2 xj, double2 vj, double rj) {
                   const double2 dxj = T.dx[j], dvj = T.dv[j];
                   D1 fx, fy;
                   if (pair_circular<D1>(d1(xi.x - xj.x, dxi.x - dxj.x),
                                         d1(xi.y - xj.y, dxi.y - dxj.y),
                                         d1(vi.x - vj.x, dvi.x - dvj.x),
                                         d1(vi.y - vj.y, dvi.y - dvj.y), D1(ri + rj), oi, fx,
                                         fy)) {
                     acc[0] += fx.d[0];
                     acc[1] += fy.d[0];
                   }
                 });
  }
  for (int mask = 1; mask < L; mask <<= 1) {
    acc[0] += __shfl_xor(acc[0], mask, 64);
    acc[1] += __shfl_xor(acc[1], mask, 64);
  }
  if (!have || q != 0) return;
  double2 df = make_double2(acc[0], acc[1]);
  if (flags & CD_USE_ADJUST) {
    // force_adjust (core/motion/adjusting.py:16-51): F = (m / tau_adj) (v0 e0 - v)
    const double s = A.mass[i] / A.tau_adj[i];
    const V2<double> e = A.e0[i];
    const double2 dv = T.dv[i];
    const double dv0 = T.dp[(size_t)i * 5 + 4];
    df.x += s * (dv0 * e.x - dv.x);
    df.y += s * (dv0 * e.y - dv.y);
  }
  T.df[i] = df;
}

__global__ void __launch_bounds__(256) k_tan_post(int n, const double *mass, double dt,
                                                  TanView T) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double c = dt / (2.0 * mass[i]);
  const double2 df = T.df[i], dg = T.dg[i], dx = T.dx[i], dv = T.dv[i];
  const double2 v = make_double2(dv.x + c * (dg.x + df.x), dv.y + c * (dg.y + df.y));
  T.dv[i] = v;
  T.dx[i] = make_double2(dx.x + dt * v.x + (dt * c) * df.x, dx.y + dt * v.y + (dt * c) * df.y);
  T.dg[i] = df;
}

}  // namespace

extern "C" {

int cd_adjoint_step(const cd_agents_t *a, const cd_step_params_t *p, const double *grad_x,
                    const cd_adjoint_t *adj, const cd_workspace_t *w, cd_stream_t stream) {
  int rc = check_step(a, p, w);
  if (rc) return rc;
  if (!adj || (a->n > 0 && (!adj->lam_x || !adj->lam_v || !adj->lam_g || !adj->lam_f ||
                            !adj->lam_vt || !adj->grad)))
    return CD_ERR_ARG;
  if (a->model != CD_CIRCULAR || a->real != CD_F64) return CD_ERR_UNSUPPORTED;
  const int allowed = CD_USE_FLUCT | CD_USE_ADJUST | CD_USE_AGENT_AGENT | CD_USE_RESET;
  if ((p->flags & ~allowed) || !(p->flags & CD_USE_RESET)) return CD_ERR_UNSUPPORTED;
  if (!(p->dt_max > 0.0) || p->dt_min != p->dt_max) return CD_ERR_UNSUPPORTED;
  if (w->n_scenarios > 0) return CD_ERR_UNSUPPORTED;
  if (a->n == 0) return CD_OK;
  hipStream_t s = (hipStream_t)stream;
  const Agents<double> A = typed<double>(a);
  const Work W = carve(w);
  const Batch B = single_batch(w, 0, p->seed, p->dt_min, p->dt_max);
  const Launch L = single_launch(w);
  AdjView J;
  J.lam_x = (double2 *)adj->lam_x;
  J.lam_v = (double2 *)adj->lam_v;
  J.lam_g = (double2 *)adj->lam_g;
  J.lam_f = (double2 *)adj->lam_f;
  J.lam_vt = (double2 *)adj->lam_vt;
  J.grad = adj->grad;
  J.grad_x = (const double2 *)grad_x;
  hipLaunchKernelGGL(k_adj_pre, dim3(grid_for(a->n, kBS)), dim3(kBS), 0, s, (int)a->n, A.mass,
                     p->dt_max, J);
  if ((rc = check_launch())) return rc;
  if (p->flags & CD_USE_AGENT_AGENT) {
    if ((rc = launch_reduce(A, W, B, L.blocks_f, true, s))) return rc;
    hipLaunchKernelGGL(k_bin_insert, dim3(L.blocks_f), dim3(kBS), 0, s, (const double2 *)A.pos,
                       p->cell_size, W, B, 1);
    hipLaunchKernelGGL(k_cell_gather<double>, dim3(grid_for((int64_t)W.cap * 64, kBS)),
                       dim3(kBS), 0, s, A, W);
    if ((rc = check_launch())) return rc;
  }
  hipLaunchKernelGGL(k_adjoint, dim3(L.blocks_e), dim3(L.threads_e), 0, s, A, W, B, J,
                     (int)p->flags, L.lanes);
  return check_launch();
}

int cd_tangent_step(const cd_agents_t *a, const cd_step_params_t *p, const cd_tangent_t *tan,
                    const cd_workspace_t *w, cd_stream_t stream) {
  int rc = check_step(a, p, w);
  if (rc) return rc;
  if (!tan || (a->n > 0 && (!tan->dx || !tan->dv || !tan->dg || !tan->df || !tan->dp)))
    return CD_ERR_ARG;
  if (a->model != CD_CIRCULAR || a->real != CD_F64) return CD_ERR_UNSUPPORTED;
  const int allowed = CD_USE_FLUCT | CD_USE_ADJUST | CD_USE_AGENT_AGENT | CD_USE_RESET;
  if ((p->flags & ~allowed) || !(p->flags & CD_USE_RESET)) return CD_ERR_UNSUPPORTED;
  if (!(p->dt_max > 0.0) || p->dt_min != p->dt_max) return CD_ERR_UNSUPPORTED;
  if (w->n_scenarios > 0) return CD_ERR_UNSUPPORTED;
  if (a->n == 0) return CD_OK;
  hipStream_t s = (hipStream_t)stream;
  const Agents<double> A = typed<double>(a);
  const Work W = carve(w);
  const Batch B = single_batch(w, 0, p->seed, p->dt_min, p->dt_max);
  const Launch L = single_launch(w);
  TanView T;
  T.dx = (double2 *)tan->dx;
  T.dv = (double2 *)tan->dv;
  T.dg = (double2 *)tan->dg;
  T.df = (double2 *)tan->df;
  T.dp = tan->dp;
  if (p->flags & CD_USE_AGENT_AGENT) {
    if ((rc = launch_reduce(A, W, B, L.blocks_f, true, s))) return rc;
    hipLaunchKernelGGL(k_bin_insert, dim3(L.blocks_f), dim3(kBS), 0, s, (const double2 *)A.pos,
                       p->cell_size, W, B, 1);
    hipLaunchKernelGGL(k_cell_gather<double>, dim3(grid_for((int64_t)W.cap * 64, kBS)),
                       dim3(kBS), 0, s, A, W);
    if ((rc = check_launch())) return rc;
  }
  hipLaunchKernelGGL(k_tangent, dim3(L.blocks_e), dim3(L.threads_e), 0, s, A, W, B, T,
                     (int)p->flags, L.lanes);
  hipLaunchKernelGGL(k_tan_post, dim3(grid_for(a->n, kBS)), dim3(kBS), 0, s, (int)a->n, A.mass,
                     p->dt_max, T);
  return check_launch();
}

int cd_rollout_record(const cd_agents_t *a, const cd_step_params_t *p, const cd_workspace_t *w,
                      int64_t steps, int64_t checkpoint, double *x_out, double *ck_pos,
                      double *ck_vel, double *ck_fprev, cd_stream_t stream) {
  int rc = check_step(a, p, w);
  if (rc) return rc;
  if (a->model != CD_CIRCULAR || a->real != CD_F64) return CD_ERR_UNSUPPORTED;
  if (steps < 0 || checkpoint < 1 || !x_out || !ck_pos || !ck_vel || !ck_fprev)
    return CD_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const size_t bytes = (size_t)a->n * 16;
  auto copy = [&](void *dst, const void *src) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s) == hipSuccess;
  };
  if (!copy(x_out, a->position)) return CD_ERR_HIP;
  for (int64_t k = 0; k < steps; ++k) {
    if (k % checkpoint == 0) {
      const size_t c = (size_t)(k / checkpoint) * a->n * 2;
      if (!copy(ck_pos + c, a->position) || !copy(ck_vel + c, a->velocity) ||
          !copy(ck_fprev + c, a->force_prev))
        return CD_ERR_HIP;
    }
    if ((rc = cd_step_tape(a, p, w, x_out + (size_t)(k + 1) * a->n * 2, nullptr, stream)))
      return rc;
  }
  return CD_OK;
}

int cd_adjoint_segment(const cd_agents_t *a, const cd_step_params_t *p, const cd_workspace_t *w,
                       int64_t k0, int64_t k1, const double *ck_pos, const double *ck_vel,
                       const double *ck_fprev, double *tape_x, double *tape_v,
                       const double *grad_x, const cd_adjoint_t *adj, cd_stream_t stream) {
  int rc = check_step(a, p, w);
  if (rc) return rc;
  if (a->model != CD_CIRCULAR || a->real != CD_F64) return CD_ERR_UNSUPPORTED;
  if (k0 < 0 || k1 <= k0 || !ck_pos || !ck_vel || !ck_fprev || !tape_x || !tape_v || !adj)
    return CD_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const size_t n2 = (size_t)a->n * 2, bytes = n2 * 8;
  auto copy = [&](void *dst, const void *src) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s) == hipSuccess;
  };
  // restore the checkpoint state (force is zero after Reset)
  if (!copy(a->position, ck_pos) || !copy(a->velocity, ck_vel) ||
      !copy(a->force_prev, ck_fprev))
    return CD_ERR_HIP;
  if ((rc = cd_reset(a, stream))) return rc;
  if ((rc = cd_step_prepare(a, p, w, (uint64_t)k0, stream))) return rc;
  // the states k0 .. k1 - 1 of the segment
  if (!copy(tape_x, a->position) || !copy(tape_v, a->velocity)) return CD_ERR_HIP;
  for (int64_t k = k0; k < k1 - 1; ++k) {
    if ((rc = cd_step_tape(a, p, w, tape_x + (size_t)(k + 1 - k0) * n2,
                           tape_v + (size_t)(k + 1 - k0) * n2, stream)))
      return rc;
  }
  // reverse sweep over the segment
  cd_agents_t view = *a;
  for (int64_t k = k1 - 1; k >= k0; --k) {
    view.position = tape_x + (size_t)(k - k0) * n2;
    view.velocity = tape_v + (size_t)(k - k0) * n2;
    if ((rc = cd_adjoint_step(&view, p, grad_x ? grad_x + (size_t)k * n2 : nullptr, adj, w,
                              stream)))
      return rc;
  }
  return CD_OK;
}

}  // extern "C"
