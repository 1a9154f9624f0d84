// Adjoint of the fused step (circular agents, fp64): the reverse sweep that
// differentiates a trajectory loss with respect to the agents' tau_0, mu,
// kappa, damping and target_velocity, and the matching forward-mode
// (tangent) step.
//
// Forward step k in the reference's order (core/integrator.py:164-176,
// crowddynamics/simulation/logic.py:59-165; G is force_prev):
//   e_k     = e0 looked up at x_k (navigation, when on; else the agent's e0)
//   F_k     = adjust(v_k, e_k; v0) + sum_j f(x_i - x_j, v_i - v_j; p_i)
//             + sum_w f_w(x_i, v_i; p_i) + noise_k
//   dt_k    = adaptive_timestep(v_k, v0) (core/integrator.py:29-97)
//   v_{k+1} = v_k + (G_k + F_k) / (2 m) * dt_k
//   x_{k+1} = x_k + v_{k+1} * dt_k + F_k / (2 m) * dt_k^2
//   G_{k+1} = F_k
// Given the adjoints (lx, lv, lg) of step k+1 the sweep forms
//   lvt = lv + dt lx,   lF = lg + dt/(2m) lvt + dt^2/(2m) lx,
//   lg_k = dt/(2m) lvt,   lv_k = lvt + (dF/dv)^T lF,   lx_k = lx + (dF/dx)^T lF,
//   grad_p += (dF/dp)^T lF,
// and the caller adds dL/dx_k.  The noise does not depend on the state or
// the parameters; Reset (logic.py:59-64) makes F_k the only force term.
// With adaptive dt, dL/d(dt_k) = sum_i lx_i . (2 (x'_i - x_i) / dt - v_i)
// + lv_i . (v'_i - v_i) / dt (x', v': state k+1, the forward step's
// derivatives dx'/d(dt) and dv'/d(dt)); while dt_k lies between its clamps
// it reaches the agent of the largest speed (d dt / d v_a = -dt v_a /
// v_max^2) and the agent of the largest target_velocity (d dt / d v0_b =
// dt / v0_max).
//
// Pair terms are owner-computes like the forward force kernel: agent i
// collects, for every neighbour j, J(d; p_i)^T lF_i from its own side and
// -J(d; p_j)^T lF_j from j's side (the pair function is odd in
// d = (x_i - x_j, v_i - v_j), so j's side is -f(d; p_j) and its Jacobian
// with respect to x_i is -J(d; p_j)).  The 2x4 Jacobian J and the
// parameter derivatives come from evaluating the forward pair function on
// dual numbers, which keeps every branch and cut-off of the step (sight,
// tau validity, force truncation, contact); so do the wall contacts
// (distance_circle_line and force_contact on dual numbers, the branch of the
// closest point taken from the forward's values) and the bilinear
// navigation direction.  Only the dt sums use atomics, on an exact 128-bit
// fixed-point pair; each agent writes only its own adjoint and gradient rows.

#include "cd_common.h"
#include "cd_dual.h"

namespace {

struct AdjView {
  double2 *lam_x, *lam_v, *lam_g;  // adjoint state, [n]
  double2 *lam_f, *lam_vt;         // lF and lvt of the current step, [n]
  double *grad;                    // [n][5]: tau_0, mu, kappa, damping, target_velocity
  const double2 *grad_x;           // dL/dx_k or null
  const double2 *x_next, *v_next;  // state k + 1 (adaptive dt)
  AdjScal *sc;                     // the step's dt scalars (adaptive dt), else null
};

// Adds a double to the exact 128-bit sum (lo, hi) of the adjoint scalars:
// the low word's carry goes into the high word, so concurrent additions in
// any order give the same pair.
__device__ __forceinline__ void atomic_add_fix(AdjScal *sc, double x) {
  Fix128 f;
  f.add(x);
  if (f.bad) atomicOr(&sc->bad, 1);
  const unsigned long long old = atomicAdd(&sc->ldt_lo, f.lo);
  const unsigned long long carry = (old + f.lo < old) ? 1ull : 0ull;
  atomicAdd((unsigned long long *)&sc->ldt_hi, (unsigned long long)f.hi + carry);
}

// Block sum of v in a fixed order (wave butterflies, then the waves in
// order on thread 0).
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double sh[16];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)((blockDim.x + 63) >> 6); ++w) t += sh[w];
  return t;  // valid on thread 0
}

// The step's dt and its maxima from the reduction of state k (the
// partials k_reduce wrote); one block.  adaptive_timestep,
// core/integrator.py:29-97.
__global__ void __launch_bounds__(256) k_adj_dt(Work W, Batch B, int adaptive) {
  const ScnDesc d = desc_t<true>(B, 0);  // one scenario (single_batch)
  const Red r = reduce_partials(W, d);
  if (threadIdx.x != 0) return;
  AdjScal *sc = adj_scal(W);
  const double vmax = red_vmax(r.v), v0max = from_ord_bits(r.v[R_V0MAX]);
  sc->dt = adaptive_dt(r.v, d.dt_min, d.dt_max);
  sc->vmax = vmax;
  sc->v0max = v0max;
  // dt = 1.1 v0max dt_max / vmax between the clamps (integrator.py:88-97)
  const double raw = (1.1 * v0max * d.dt_max) / vmax;
  sc->inside = (adaptive && d.dt_min < d.dt_max && vmax > 0.0 && !(raw > d.dt_max) &&
                !(raw < d.dt_min))
                   ? 1
                   : 0;
  sc->arg_v = 0x7fffffff;
  sc->arg_v0 = 0x7fffffff;
  sc->bad = 0;
  sc->ldt_lo = 0ull;
  sc->ldt_hi = 0;
  sc->ddt = 0.0;
}

// The agents attaining the maxima of speed and target_velocity (the
// lowest index of each), for dt's derivative.
__device__ __forceinline__ void note_maxima(AdjScal *sc, int i, double2 v, double v0) {
  if (dhypot(v.x, v.y) == sc->vmax) atomicMin(&sc->arg_v, i);
  if (v0 == sc->v0max) atomicMin(&sc->arg_v0, i);
}

__global__ void __launch_bounds__(256) k_adj_pre(int n, const double *mass, const double *v0,
                                                 double dt_fixed, const double2 *pos,
                                                 const V2<double> *vel, AdjView J) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool adaptive = J.sc != nullptr;  // uniform
  const double dt = adaptive ? J.sc->dt : dt_fixed;
  double ldt = 0.0;
  if (i < n) {
    const double c = dt / (2.0 * mass[i]);
    const double2 lx = J.lam_x[i], lv = J.lam_v[i], lg = J.lam_g[i];
    const double2 lvt = make_double2(lv.x + dt * lx.x, lv.y + dt * lx.y);
    const double c2 = dt * c;
    J.lam_f[i] = make_double2(lg.x + c * lvt.x + c2 * lx.x, lg.y + c * lvt.y + c2 * lx.y);
    J.lam_vt[i] = lvt;
    J.lam_g[i] = make_double2(c * lvt.x, c * lvt.y);
    if (adaptive && J.sc->inside) {
      const double2 x = pos[i], xn = J.x_next[i], vn = J.v_next[i];
      const V2<double> vv = vel[i];
      const double2 v = make_double2(vv.x, vv.y);
      ldt = lx.x * (2.0 * (xn.x - x.x) / dt - v.x) + lx.y * (2.0 * (xn.y - x.y) / dt - v.y) +
            lv.x * ((vn.x - v.x) / dt) + lv.y * ((vn.y - v.y) / dt);
      note_maxima(J.sc, i, v, v0[i]);
    }
  }
  if (adaptive) {  // uniform over the block
    const double t = block_sum(ldt);
    if (threadIdx.x == 0 && J.sc->inside) atomic_add_fix(J.sc, t);
  }
}

// After the adjoint kernel: dt's derivative joins the adjoint of the
// fastest agent's velocity and the gradient of the largest target_velocity.
__global__ void k_adj_dt_apply(AdjScal *sc, const V2<double> *vel, AdjView J, int n) {
  if (!sc->inside) return;
  Fix128 f;
  f.hi = sc->ldt_hi;
  f.lo = sc->ldt_lo;
  f.bad = sc->bad != 0;
  const double ldt = f.value(), dt = sc->dt;
  const int a = sc->arg_v, b = sc->arg_v0;
  if (a < n) {
    const V2<double> v = vel[a];
    const double s = -ldt * dt / (sc->vmax * sc->vmax);
    double2 lv = J.lam_v[a];
    lv.x += s * v.x;
    lv.y += s * v.y;
    J.lam_v[a] = lv;
  }
  if (b < n) J.grad[(size_t)b * 5 + 4] += ldt * dt / sc->v0max;
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

// Contact force of a circular agent at (x, y) with velocity (vx, vy) and
// radius r on wall segment s with frame fr (seg_frame), on dual numbers:
// distance_circle_line (core/distance.py:108-146) with the branch of the
// closest point decided on the values, which are those of the forward's
// double evaluation (the same operations), then force_contact with the wall
// at rest (core/interactions.py:107-141).  False when out of contact.
template <int K>
__device__ __forceinline__ bool wall_contact_dual(const Dual<K> &x, const Dual<K> &y,
                                                  const Dual<K> &vx, const Dual<K> &vy, double r,
                                                  const double4 &s, const double4 &fr,
                                                  const Dual<K> &mu, const Dual<K> &kappa,
                                                  const Dual<K> &damping, Dual<K> &fx,
                                                  Dual<K> &fy) {
  using D = Dual<K>;
  const D tx(fr.x), ty(fr.y);
  const double l_w = fr.z;
  const D q0x = x - D(s.x), q0y = y - D(s.y);
  const D q1x = x - D(s.z), q1y = y - D(s.w);
  const D l_t = -(tx * q1x + ty * q1y) - (tx * q0x + ty * q0y);
  D d_iw, nx, ny;
  if (l_t.v > l_w) {
    d_iw = dhypot(q0x, q0y);
    nx = q0x / d_iw;
    ny = q0y / d_iw;
  } else if (l_t.v < -l_w) {
    d_iw = dhypot(q1x, q1y);
    nx = q1x / d_iw;
    ny = q1y / d_iw;
  } else {
    const D l_n = (-ty) * q0x + tx * q0y;  // n_w = rotate90(t)
    const double sg = np_sign(l_n.v);
    d_iw = D(sg) * l_n;  // |l_n|, derivative sign(l_n) dl_n
    nx = D(sg * (-fr.y));
    ny = D(sg * fr.x);
  }
  const D h = d_iw - D(r);
  if (!(h.v < 0.0)) return false;
  force_contact<D>(h, nx, ny, vx, vy, ny, -nx, mu, kappa, damping, fx, fy);
  return true;
}

// Calls fn(segment, frame) for the wall segments of scenario d in contact
// with agent i's disc at xi (the forward's double distance), lane q of L
// taking segments q, q + L, ...; a segment whose bounding box lies beyond
// the radius (+ 1e-6 m) is out of contact, as in the force kernel.
template <class Fn>
__device__ __forceinline__ void wall_contacts(const FusedArgs &fa, const ScnDesc &d, double2 xi,
                                              double ri, int q, int L, Fn &&fn) {
  const double reach = ri + 1e-6, reach2 = reach * reach;
  for (int w = d.w0 + q; w < d.w1; w += L) {
    const double4 sg = fa.segs[w];
    const double bx = fmax(fmax(fmin(sg.x, sg.z) - xi.x, xi.x - fmax(sg.x, sg.z)), 0.0);
    const double by = fmax(fmax(fmin(sg.y, sg.w) - xi.y, xi.y - fmax(sg.y, sg.w)), 0.0);
    if (bx * bx + by * by > reach2) continue;
    const double4 fr = seg_frame(sg.x, sg.y, sg.z, sg.w);
    double h, nx, ny;
    distance_circle_line(xi.x, xi.y, ri, sg.x, sg.y, sg.z, sg.w, fr, h, nx, ny);
    if (h < 0.0) fn(sg, fr);
  }
}

// The navigation direction at (px, py) (nav_fetch / nav_finish, the
// forward's values) and its Jacobian J = d(ex, ey)/d(px, py) (row-major):
// in the bilinear mode the gradient of the bilinear potential on dual
// numbers (the cell fixed, d fx / d px = d fy / d py = 1 / step), at the
// nearest node zero.  False when the lookup keeps the agent's e0
// (core/steering/navigation.py:59-78).
__device__ __forceinline__ bool nav_dual(const FusedArgs &fa, int64_t target, double2 p,
                                         double &ex, double &ey, double J[4]) {
  const NavFetch nf = nav_fetch(p.x, p.y, target, fa.maps, fa.nav_T, fa.nav_rows, fa.nav_cols,
                                fa.nav_minx, fa.nav_miny, fa.nav_step, fa.nav_mode);
  J[0] = J[1] = J[2] = J[3] = 0.0;
  if (!nav_finish(nf, ex, ey)) return false;
  if (nf.mode == 1) {
    using D2 = Dual<2>;
    D2 fx(nf.fx), fy(nf.fy);
    fx.d[0] = 1.0 / nf.step;
    fy.d[1] = 1.0 / nf.step;
    const D2 one(1.0), st(nf.step);
    const D2 gx = ((D2(nf.p01) - D2(nf.p00)) * (one - fy) + (D2(nf.p11) - D2(nf.p10)) * fy) / st;
    const D2 gy = ((D2(nf.p10) - D2(nf.p00)) * (one - fx) + (D2(nf.p11) - D2(nf.p01)) * fx) / st;
    const D2 g = dhypot(gx, gy);
    const D2 exd = -gx / g, eyd = -gy / g;
    J[0] = exd.d[0];
    J[1] = exd.d[1];
    J[2] = eyd.d[0];
    J[3] = eyd.d[1];
  }
  return true;
}

// Agent i's target direction of step k and its Jacobian (zero when
// navigation is off or the lookup keeps e0, which then comes from the
// agents' e0 and raises CD_STATUS_NAV_KEPT).
__device__ __forceinline__ V2<double> step_e0(const Agents<double> &A, const FusedArgs &fa,
                                              const Work &W, int i, double2 xi, double J[4]) {
  J[0] = J[1] = J[2] = J[3] = 0.0;
  if (fa.flags & CD_USE_NAV) {
    double ex, ey;
    if (nav_dual(fa, A.target[i], xi, ex, ey, J)) {
      V2<double> e;
      e.x = ex;
      e.y = ey;
      return e;
    }
    atomicOr((unsigned long long *)&W.st->status, (unsigned long long)CD_STATUS_NAV_KEPT);
  }
  return A.e0[i];
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
                                                 FusedArgs fa, int L) {
  const int flags = fa.flags;
  const int2 sk = block_e_t<true>(B);  // one scenario (single_batch)
  const ScnDesc d = desc_t<true>(B, sk.x);
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
  if (have && (flags & CD_USE_OBSTACLE)) {
    // wall contacts, f_w(x_i, v_i; mu_i, kappa_i, damping_i)
    const double2 xi = A.pos[i];
    const V2<double> v = A.vel[i];
    const double ri = A.radius[i], mu = A.mu[i], ka = A.kappa[i], da = A.damping[i];
    wall_contacts(fa, d, xi, ri, q, L, [&](const double4 &sg, const double4 &fr) {
      D4 fx, fy;
      wall_contact_dual<4>(D4::seed(xi.x, 0), D4::seed(xi.y, 1), D4::seed(v.x, 2),
                           D4::seed(v.y, 3), ri, sg, fr, D4(mu), D4(ka), D4(da), fx, fy);
#pragma unroll
      for (int k = 0; k < 4; ++k) acc[k] += lfi.x * fx.d[k] + lfi.y * fy.d[k];
      wall_contact_dual<4>(D4(xi.x), D4(xi.y), D4(v.x), D4(v.y), ri, sg, fr, D4::seed(mu, 1),
                           D4::seed(ka, 2), D4::seed(da, 3), fx, fy);
#pragma unroll
      for (int k = 1; k < 4; ++k) acc[4 + k] += lfi.x * fx.d[k] + lfi.y * fy.d[k];
    });
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
  double2 lx = J.lam_x[i];
  double g_v0 = 0.0;
  if (flags & CD_USE_ADJUST) {
    // force_adjust (core/motion/adjusting.py:16-51): F = (m / tau_adj) (v0 e0 - v),
    // e0 = e_k(x_k) with navigation (logic.py:149-165)
    const double s = A.mass[i] / A.tau_adj[i];
    double Je[4];
    const V2<double> e = step_e0(A, fa, W, i, A.pos[i], Je);
    lv.x -= s * lfi.x;
    lv.y -= s * lfi.y;
    g_v0 = s * (e.x * lfi.x + e.y * lfi.y);
    const double sv = s * A.v0[i];
    lx.x += sv * (Je[0] * lfi.x + Je[2] * lfi.y);
    lx.y += sv * (Je[1] * lfi.x + Je[3] * lfi.y);
  }
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
  double2 *dx, *dv, *dg;           // tangent state, [n]
  double2 *df;                     // dF_k of the current step, [n]
  const double *dp;                // [n][5]: tau_0, mu, kappa, damping, target_velocity
  const double2 *x_next, *v_next;  // state k + 1 (adaptive dt)
  AdjScal *sc;                     // the step's dt scalars (adaptive dt), else null
};

using D1 = Dual<1>;

__device__ __forceinline__ D1 d1(double v, double t) {
  D1 r(v);
  r.d[0] = t;
  return r;
}

__global__ void __launch_bounds__(256) k_tangent(Agents<double> A, Work W, Batch B, TanView T,
                                                 FusedArgs fa, int L) {
  const int flags = fa.flags;
  const int2 sk = block_e_t<true>(B);  // one scenario (single_batch)
  const ScnDesc d = desc_t<true>(B, sk.x);
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
  if (have && (flags & CD_USE_OBSTACLE)) {
    // wall contacts, f_w(x_i, v_i; mu_i, kappa_i, damping_i)
    const double2 xi = A.pos[i];
    const V2<double> v = A.vel[i];
    const double2 dxi = T.dx[i], dvi = T.dv[i];
    const double *dpi = T.dp + (size_t)i * 5;
    const double ri = A.radius[i];
    wall_contacts(fa, d, xi, ri, q, L, [&](const double4 &sg, const double4 &fr) {
      D1 fx, fy;
      wall_contact_dual<1>(d1(xi.x, dxi.x), d1(xi.y, dxi.y), d1(v.x, dvi.x), d1(v.y, dvi.y), ri,
                           sg, fr, d1(A.mu[i], dpi[1]), d1(A.kappa[i], dpi[2]),
                           d1(A.damping[i], dpi[3]), fx, fy);
      acc[0] += fx.d[0];
      acc[1] += fy.d[0];
    });
  }
  for (int mask = 1; mask < L; mask <<= 1) {
    acc[0] += __shfl_xor(acc[0], mask, 64);
    acc[1] += __shfl_xor(acc[1], mask, 64);
  }
  if (!have || q != 0) return;
  double2 df = make_double2(acc[0], acc[1]);
  if (flags & CD_USE_ADJUST) {
    // force_adjust (core/motion/adjusting.py:16-51): F = (m / tau_adj) (v0 e0 - v),
    // e0 = e_k(x_k) with navigation (logic.py:149-165)
    const double s = A.mass[i] / A.tau_adj[i];
    double Je[4];
    const V2<double> e = step_e0(A, fa, W, i, A.pos[i], Je);
    const double2 dv = T.dv[i], dx = T.dx[i];
    const double dv0 = T.dp[(size_t)i * 5 + 4], sv = s * A.v0[i];
    df.x += s * (dv0 * e.x - dv.x) + sv * (Je[0] * dx.x + Je[1] * dx.y);
    df.y += s * (dv0 * e.y - dv.y) + sv * (Je[2] * dx.x + Je[3] * dx.y);
  }
  T.df[i] = df;
  if (T.sc && T.sc->inside) {
    const V2<double> v = A.vel[i];
    note_maxima(T.sc, i, make_double2(v.x, v.y), A.v0[i]);
  }
}

// The tangent of dt: d dt = -dt (v_a . dv_a) / v_max^2 + dt dv0_b / v0_max
// for the agents a, b of the largest speed and target_velocity, while dt
// lies between its clamps.
__global__ void k_tan_dt(AdjScal *sc, const V2<double> *vel, TanView T, int n) {
  double ddt = 0.0;
  if (sc->inside) {
    const double dt = sc->dt;
    const int a = sc->arg_v, b = sc->arg_v0;
    if (a < n) {
      const V2<double> v = vel[a];
      const double2 dv = T.dv[a];
      ddt -= dt * (v.x * dv.x + v.y * dv.y) / (sc->vmax * sc->vmax);
    }
    if (b < n) ddt += dt * T.dp[(size_t)b * 5 + 4] / sc->v0max;
  }
  sc->ddt = ddt;
}

__global__ void __launch_bounds__(256) k_tan_post(int n, const double *mass, double dt_fixed,
                                                  const double2 *pos, const V2<double> *vel,
                                                  TanView T) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool adaptive = T.sc != nullptr;
  const double dt = adaptive ? T.sc->dt : dt_fixed;
  const double c = dt / (2.0 * mass[i]);
  const double2 df = T.df[i], dg = T.dg[i], dx = T.dx[i], dv = T.dv[i];
  double2 v = make_double2(dv.x + c * (dg.x + df.x), dv.y + c * (dg.y + df.y));
  double2 x = make_double2(dx.x + dt * v.x + (dt * c) * df.x, dx.y + dt * v.y + (dt * c) * df.y);
  if (adaptive && T.sc->ddt != 0.0) {
    // the step's derivatives dv'/d(dt) = (v' - v) / dt and, beyond the
    // dv' dt term, dx'/d(dt) = 2 (x' - x) / dt - v'
    const double ddt = T.sc->ddt;
    const double2 x0 = pos[i], xn = T.x_next[i], vn = T.v_next[i];
    const V2<double> v0 = vel[i];
    const double2 av = make_double2((vn.x - v0.x) / dt * ddt, (vn.y - v0.y) / dt * ddt);
    v.x += av.x;
    v.y += av.y;
    x.x += dt * av.x + (2.0 * (xn.x - x0.x) / dt - vn.x) * ddt;
    x.y += dt * av.y + (2.0 * (xn.y - x0.y) / dt - vn.y) * ddt;
  }
  T.dv[i] = v;
  T.dx[i] = x;
  T.dg[i] = df;
}

// The adjoint and tangent steps' scope: circular agents in fp64, a single
// scenario, RESET set, any other node of the standard tree; adaptive dt
// needs the state of step k + 1.
int check_diff_scope(const cd_agents_t *a, const cd_step_params_t *p, const cd_workspace_t *w,
                     const double *x_next, const double *v_next) {
  if (a->model != CD_CIRCULAR || a->real != CD_F64) return CD_ERR_UNSUPPORTED;
  const int allowed = CD_USE_FLUCT | CD_USE_NAV | CD_USE_ADJUST | CD_USE_AGENT_AGENT |
                      CD_USE_OBSTACLE | CD_USE_RESET;
  if ((p->flags & ~allowed) || !(p->flags & CD_USE_RESET)) return CD_ERR_UNSUPPORTED;
  if (!(p->dt_max > 0.0) || !(p->dt_min > 0.0) || p->dt_min > p->dt_max) return CD_ERR_ARG;
  if (p->dt_min < p->dt_max && a->n > 0 && (!x_next || !v_next)) return CD_ERR_ARG;
  if (w->n_scenarios > 0) return CD_ERR_UNSUPPORTED;
  return CD_OK;
}

// Binning of state k (pairs) and the reduction adaptive dt is formed from.
int diff_prepare(const Agents<double> &A, const cd_step_params_t *p, const Work &W,
                 const Batch &B, const Launch &L, bool adaptive, hipStream_t s) {
  int rc;
  const bool pairs = (p->flags & CD_USE_AGENT_AGENT) != 0;
  if (pairs || adaptive)
    if ((rc = launch_reduce(A, W, B, L.blocks_f, true, s))) return rc;
  if (adaptive) hipLaunchKernelGGL(k_adj_dt, dim3(1), dim3(kBS), 0, s, W, B, 1);
  if (pairs) {
    hipLaunchKernelGGL(k_bin_insert, dim3(L.blocks_f), dim3(kBS), 0, s, (const double2 *)A.pos,
                       p->cell_size, W, B, 1);
    hipLaunchKernelGGL(k_cell_gather<double>, dim3(grid_for((int64_t)W.cap * 64, kBS)),
                       dim3(kBS), 0, s, A, W);
  }
  return check_launch();
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
  if ((rc = check_diff_scope(a, p, w, adj->x_next, adj->v_next))) return rc;
  if (a->n == 0) return CD_OK;
  hipStream_t s = (hipStream_t)stream;
  const Agents<double> A = typed<double>(a);
  const Work W = carve(w);
  const Batch B = single_batch(w, (p->flags & CD_USE_OBSTACLE) ? p->n_segments : 0, p->seed,
                               p->dt_min, p->dt_max);
  const Launch L = single_launch(w);
  const bool adaptive = p->dt_min < p->dt_max;
  AdjView J;
  J.lam_x = (double2 *)adj->lam_x;
  J.lam_v = (double2 *)adj->lam_v;
  J.lam_g = (double2 *)adj->lam_g;
  J.lam_f = (double2 *)adj->lam_f;
  J.lam_vt = (double2 *)adj->lam_vt;
  J.grad = adj->grad;
  J.grad_x = (const double2 *)grad_x;
  J.x_next = (const double2 *)adj->x_next;
  J.v_next = (const double2 *)adj->v_next;
  J.sc = adaptive ? adj_scal(W) : nullptr;
  if ((rc = diff_prepare(A, p, W, B, L, adaptive, s))) return rc;
  hipLaunchKernelGGL(k_adj_pre, dim3(grid_for(a->n, kBS)), dim3(kBS), 0, s, (int)a->n, A.mass,
                     (const double *)A.v0, p->dt_max, (const double2 *)A.pos,
                     (const V2<double> *)A.vel, J);
  hipLaunchKernelGGL(k_adjoint, dim3(L.blocks_e), dim3(L.threads_e), 0, s, A, W, B, J,
                     fused_args(p), L.lanes);
  if (adaptive)
    hipLaunchKernelGGL(k_adj_dt_apply, dim3(1), dim3(1), 0, s, J.sc, (const V2<double> *)A.vel,
                       J, (int)a->n);
  return check_launch();
}

int cd_tangent_step(const cd_agents_t *a, const cd_step_params_t *p, const cd_tangent_t *tan,
                    const cd_workspace_t *w, cd_stream_t stream) {
  int rc = check_step(a, p, w);
  if (rc) return rc;
  if (!tan || (a->n > 0 && (!tan->dx || !tan->dv || !tan->dg || !tan->df || !tan->dp)))
    return CD_ERR_ARG;
  if ((rc = check_diff_scope(a, p, w, tan->x_next, tan->v_next))) return rc;
  if (a->n == 0) return CD_OK;
  hipStream_t s = (hipStream_t)stream;
  const Agents<double> A = typed<double>(a);
  const Work W = carve(w);
  const Batch B = single_batch(w, (p->flags & CD_USE_OBSTACLE) ? p->n_segments : 0, p->seed,
                               p->dt_min, p->dt_max);
  const Launch L = single_launch(w);
  const bool adaptive = p->dt_min < p->dt_max;
  TanView T;
  T.dx = (double2 *)tan->dx;
  T.dv = (double2 *)tan->dv;
  T.dg = (double2 *)tan->dg;
  T.df = (double2 *)tan->df;
  T.dp = tan->dp;
  T.x_next = (const double2 *)tan->x_next;
  T.v_next = (const double2 *)tan->v_next;
  T.sc = adaptive ? adj_scal(W) : nullptr;
  if ((rc = diff_prepare(A, p, W, B, L, adaptive, s))) return rc;
  hipLaunchKernelGGL(k_tangent, dim3(L.blocks_e), dim3(L.threads_e), 0, s, A, W, B, T,
                     fused_args(p), L.lanes);
  if (adaptive)
    hipLaunchKernelGGL(k_tan_dt, dim3(1), dim3(1), 0, s, T.sc, (const V2<double> *)A.vel, T,
                       (int)a->n);
  hipLaunchKernelGGL(k_tan_post, dim3(grid_for(a->n, kBS)), dim3(kBS), 0, s, (int)a->n, A.mass,
                     p->dt_max, (const double2 *)A.pos, (const V2<double> *)A.vel, T);
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
  // the states k0 .. k1 - 1 of the segment (and k1 with adaptive dt, whose
  // derivative needs the state after each step)
  const bool adaptive = p->dt_min < p->dt_max;
  const int64_t last = adaptive ? k1 : k1 - 1;
  if (!copy(tape_x, a->position) || !copy(tape_v, a->velocity)) return CD_ERR_HIP;
  for (int64_t k = k0; k < last; ++k) {
    if ((rc = cd_step_tape(a, p, w, tape_x + (size_t)(k + 1 - k0) * n2,
                           tape_v + (size_t)(k + 1 - k0) * n2, stream)))
      return rc;
  }
  // reverse sweep over the segment
  cd_agents_t view = *a;
  cd_adjoint_t step_adj = *adj;
  for (int64_t k = k1 - 1; k >= k0; --k) {
    view.position = tape_x + (size_t)(k - k0) * n2;
    view.velocity = tape_v + (size_t)(k - k0) * n2;
    step_adj.x_next = adaptive ? tape_x + (size_t)(k + 1 - k0) * n2 : nullptr;
    step_adj.v_next = adaptive ? tape_v + (size_t)(k + 1 - k0) * n2 : nullptr;
    if ((rc = cd_adjoint_step(&view, p, grad_x ? grad_x + (size_t)k * n2 : nullptr, &step_adj,
                              w, stream)))
      return rc;
  }
  return CD_OK;
}

}  // extern "C"
