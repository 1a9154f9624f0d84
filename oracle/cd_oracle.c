/*
 * cd_oracle.c -- CPU restatement (fp64, strict IEEE, single thread) of the
 * crowddynamics per-timestep agent update.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the reference the GPU path is
 * compared against and the host-CPU baseline that bench.py times.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product path (crowddynamics_amd) never links, loads or falls back to it.
 *
 * Every function names the reference function it restates (file:line, paths
 * relative to the reference repository root).  The operation order of each
 * floating-point expression follows the reference source, and the file must be
 * built with -ffp-contract=off (numba does not contract a*b+c into an FMA).
 *
 * Binning contract (the reference delegates binning to the un-vendored
 * `cell_lists` package, so this contract is the build's own definition):
 *   origin   o = component-wise minimum of all positions
 *   cell     l_i = (trunc((x_i - o_x) / c), trunc((y_i - o_y) / c))
 *   shape    (nx, ny) = (max l_x + 1, max l_y + 1)
 *   flat     cell(l) = l_x * ny + l_y
 *   order    points sorted stably by flat cell (ties keep agent index order)
 *   pairs    every unordered {i, j}, i != j, with |dl_x| <= 1 and |dl_y| <= 1,
 *            evaluated once as interaction(min(i,j), max(i,j)).
 *   iteration order: cells in flat order; within a cell all (a<b) slot pairs,
 *            then the half stencil (0,+1), (+1,-1), (+1,0), (+1,+1).
 *
 * Parity status of each part is listed in the header of oracle/oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* When set (only inside orc_step_omp), the per-agent loops below run over
 * OpenMP threads; each agent's arithmetic and summation order is unchanged. */
static int g_par = 0;

#define ORC_SIGHT_SOC 3.0   /* core/interactions.py:45 */
#define ORC_F_SOC_MAX 2e3   /* core/motion/power_law.py:52 */
#define ORC_TAU_MAX 30.0    /* core/motion/power_law.py:243 */

typedef struct orc_agents {
  int64_t n;
  int32_t model; /* 0 = circular, 1 = three_circle */
  int32_t _pad;
  /* vectors, interleaved [n][2] */
  double *position, *velocity, *target_direction, *force, *force_prev;
  double *position_ls, *position_rs;
  /* scalars [n] */
  double *radius, *r_t, *r_s, *r_ts, *mass, *inertia_rot;
  double *target_velocity, *target_angular_velocity;
  double *tau_adj, *k_soc, *tau_0, *mu, *kappa, *damping, *std_rand_force;
  double *orientation, *angular_velocity, *target_orientation, *torque,
      *torque_prev, *tau_rot, *std_rand_torque;
  int64_t *target;
} orc_agents;

/* ------------------------------------------------------------------ */
/* vector2D (core/vector2D.py)                                          */
/* ------------------------------------------------------------------ */

static inline double dot2(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1]; /* vector2D.py:128 */
}
static inline double cross2(const double *a, const double *b) {
  return a[0] * b[1] - a[1] * b[0]; /* vector2D.py:149 */
}
static inline double length2(const double *v) {
  return hypot(v[0], v[1]); /* vector2D.py:113 (np.hypot) */
}

/* Python float remainder a % b (b > 0), as numba implements it. */
static double py_fmod(double a, double b) {
  double m = fmod(a, b);
  if (m != 0.0) {
    if ((b < 0.0) != (m < 0.0)) m += b;
  } else {
    m = copysign(0.0, b);
  }
  return m;
}

/* wrap_to_pi, core/vector2D.py:8-36 */
double orc_wrap_to_pi(double rad) {
  const double two_pi = 2.0 * M_PI;
  double r = py_fmod(rad, two_pi);
  if (rad < 0.0 && r == M_PI) return -M_PI;
  if (r > M_PI) return r - two_pi;
  return r;
}

/* truncate, core/vector2D.py:167-187 */
static inline void truncate2(double *v, double l) {
  double vlen = length2(v);
  if (vlen > l) {
    double s = l / vlen;
    v[0] *= s;
    v[1] *= s;
  }
}

/* ------------------------------------------------------------------ */
/* distance (core/distance.py)                                          */
/* ------------------------------------------------------------------ */

/* distance_circles, core/distance.py:16-47 */
static inline void distance_circles(const double *x0, double r0,
                                    const double *x1, double r1, double *h,
                                    double *n) {
  double x[2] = {x0[0] - x1[0], x0[1] - x1[1]};
  double d = length2(x);
  double r_tot = r0 + r1;
  *h = d - r_tot;
  if (d == 0.0) {
    n[0] = 0.0;
    n[1] = 0.0;
  } else {
    n[0] = x[0] / d;
    n[1] = x[1] / d;
  }
}

/* distance_three_circles, core/distance.py:50-105 (r_moment1 keeps the
 * reference's x0[j_min] at :103). */
static void distance_three_circles(const double *x0[3], const double r0[3],
                                   const double *x1[3], const double r1[3],
                                   double *h_out, double *normal,
                                   double *rm0, double *rm1) {
  double h_min = NAN;
  normal[0] = 0.0;
  normal[1] = 0.0;
  int i_min = 0, j_min = 0;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) {
      double h, n[2];
      distance_circles(x0[i], r0[i], x1[j], r1[j], &h, n);
      if (h < h_min || isnan(h_min)) {
        h_min = h;
        normal[0] = n[0];
        normal[1] = n[1];
        i_min = i;
        j_min = j;
      }
    }
  }
  for (int k = 0; k < 2; ++k) {
    rm0[k] = (x0[i_min][k] + r0[i_min] * normal[k]) - x0[0][k];
    rm1[k] = (x0[j_min][k] - r1[j_min] * normal[k]) - x1[0][k];
  }
  *h_out = h_min;
}

static inline double np_sign(double x) {
  if (x > 0.0) return 1.0;
  if (x < 0.0) return -1.0;
  if (x == 0.0) return 0.0;
  return x; /* NaN */
}

/* distance_circle_line, core/distance.py:108-146 */
void orc_distance_circle_line(const double *x, double r, const double *p0,
                              const double *p1, double *h_out, double *n_iw) {
  double d[2] = {p1[0] - p0[0], p1[1] - p0[1]};
  double l_w = length2(d);
  double t_w[2] = {d[0] / l_w, d[1] / l_w};
  double n_w[2] = {-t_w[1], t_w[0]}; /* rotate90 */
  double q0[2] = {x[0] - p0[0], x[1] - p0[1]};
  double q1[2] = {x[0] - p1[0], x[1] - p1[1]};
  double l_t = -dot2(t_w, q1) - dot2(t_w, q0);
  double d_iw;
  if (l_t > l_w) {
    d_iw = length2(q0);
    n_iw[0] = q0[0] / d_iw;
    n_iw[1] = q0[1] / d_iw;
  } else if (l_t < -l_w) {
    d_iw = length2(q1);
    n_iw[0] = q1[0] / d_iw;
    n_iw[1] = q1[1] / d_iw;
  } else {
    double l_n = dot2(n_w, q0);
    d_iw = fabs(l_n);
    double s = np_sign(l_n);
    n_iw[0] = s * n_w[0];
    n_iw[1] = s * n_w[1];
  }
  *h_out = d_iw - r;
}

/* distance_three_circle_line, core/distance.py:149-180 */
static void distance_three_circle_line(const double *x[3], const double r[3],
                                       const double *p0, const double *p1,
                                       double *h_out, double *normal,
                                       double *rm) {
  double h_min = NAN;
  normal[0] = 0.0;
  normal[1] = 0.0;
  int i_min = 0;
  for (int i = 0; i < 3; ++i) {
    double h, n[2];
    orc_distance_circle_line(x[i], r[i], p0, p1, &h, n);
    if (h < h_min || isnan(h_min)) {
      h_min = h;
      normal[0] = n[0];
      normal[1] = n[1];
      i_min = i;
    }
  }
  for (int k = 0; k < 2; ++k)
    rm[k] = (x[i_min][k] - r[i_min] * normal[k]) - x[0][k];
  *h_out = h_min;
}

/* Test access to distance_three_circles: x0, x1 are [3][2] part centres. */
void orc_distance_three_circles(const double *x0, const double *r0, const double *x1,
                                const double *r1, double *h, double *n, double *rm0,
                                double *rm1) {
  const double *a[3] = {x0, x0 + 2, x0 + 4}, *b[3] = {x1, x1 + 2, x1 + 4};
  distance_three_circles(a, r0, b, r1, h, n, rm0, rm1);
}

/* ------------------------------------------------------------------ */
/* motion (core/motion/)                                                */
/* ------------------------------------------------------------------ */

/* magnitude, core/motion/power_law.py:83-102 */
static inline double magnitude(double tau, double tau_0) {
  return ((2.0 / tau + 1.0 / tau_0) * exp(-tau / tau_0)) / (tau * tau);
}

/* force_contact, core/motion/contact.py:12-48 */
static inline void force_contact(double h, const double *n, const double *v,
                                 const double *t, double mu, double kappa,
                                 double damping, double *out) {
  double kvt = kappa * dot2(v, t);
  double cvn = damping * dot2(v, n);
  for (int k = 0; k < 2; ++k)
    out[k] = (-h) * (mu * n[k] - kvt * t[k]) + cvn * n[k];
}

/* force_social_circular, core/motion/power_law.py:213-259 */
static void force_social_circular(const orc_agents *A, int64_t i, int64_t j,
                                  double *fi, double *fj) {
  fi[0] = fi[1] = fj[0] = fj[1] = 0.0;
  const double *xi = A->position + 2 * i, *xj = A->position + 2 * j;
  const double *vi = A->velocity + 2 * i, *vj = A->velocity + 2 * j;
  double x_rel[2] = {xi[0] - xj[0], xi[1] - xj[1]};
  double v_rel[2] = {vi[0] - vj[0], vi[1] - vj[1]};
  double r_tot = A->radius[i] + A->radius[j];
  double a = dot2(v_rel, v_rel);
  double b = -dot2(x_rel, v_rel);
  double c = dot2(x_rel, x_rel) - r_tot * r_tot;
  double d = sqrt(b * b - a * c);
  if (isnan(d) || d == 0.0 || a == 0.0) return;
  double tau = (b - d) / a;
  if (tau <= 0.0 || tau > ORC_TAU_MAX) return;
  double grad[2];
  for (int k = 0; k < 2; ++k) grad[k] = (v_rel[k] - (v_rel[k] * b + x_rel[k] * a) / d) / a;
  double si = (-A->mass[i]) * A->k_soc[i];
  double sj = (-A->mass[j]) * A->k_soc[j];
  double mi = magnitude(tau, A->tau_0[i]);
  double mj = magnitude(tau, A->tau_0[j]);
  for (int k = 0; k < 2; ++k) {
    fi[k] = 0.0 + (si * grad[k]) * mi;
    fj[k] = 0.0 - (sj * grad[k]) * mj;
  }
  truncate2(fi, ORC_F_SOC_MAX);
  truncate2(fj, ORC_F_SOC_MAX);
}

/* force_social_three_circle, core/motion/power_law.py:262-363 */
static void force_social_three_circle(const orc_agents *A, int64_t i,
                                      int64_t j, double *fi, double *fj) {
  fi[0] = fi[1] = fj[0] = fj[1] = 0.0;
  const double *vi = A->velocity + 2 * i, *vj = A->velocity + 2 * j;
  double v_rel[2] = {vi[0] - vj[0], vi[1] - vj[1]};
  double a = dot2(v_rel, v_rel);
  if (a == 0.0) return;
  const double *x_i[3] = {A->position + 2 * i, A->position_ls + 2 * i,
                          A->position_rs + 2 * i};
  const double *x_j[3] = {A->position + 2 * j, A->position_ls + 2 * j,
                          A->position_rs + 2 * j};
  double r_i[3] = {A->r_t[i], A->r_s[i], A->r_s[i]};
  double r_j[3] = {A->r_t[j], A->r_s[j], A->r_s[j]};
  int contact_i = 0, contact_j = 0;
  double tau = NAN, b_min = NAN, d_min = NAN;
  for (int pi = 0; pi < 3; ++pi) {
    for (int pj = 0; pj < 3; ++pj) {
      double x_rel[2] = {x_i[pi][0] - x_j[pj][0], x_i[pi][1] - x_j[pj][1]};
      double r_tot = r_i[pi] + r_j[pj];
      double b = -dot2(x_rel, v_rel);
      double c = dot2(x_rel, x_rel) - r_tot * r_tot;
      double d = sqrt(b * b - a * c);
      if (isnan(d) || d == 0.0) continue;
      double tau_new = (b - d) / a;
      if (isnan(tau) || (0.0 < tau_new && tau_new < tau)) {
        contact_i = pi;
        contact_j = pj;
        tau = tau_new;
        b_min = b;
        d_min = d;
      }
    }
  }
  if (isnan(tau) || tau <= 0.0) return;
  double r_off_i[2] = {0.0, 0.0}, r_off_j[2] = {0.0, 0.0};
  if (contact_i == 1 || contact_i == 2) {
    double phi = A->orientation[i];
    double e[2] = {A->r_ts[i] * sin(phi), A->r_ts[i] * (-cos(phi))};
    if (contact_i == 1) {
      r_off_i[0] += e[0];
      r_off_i[1] += e[1];
    } else {
      r_off_i[0] -= e[0];
      r_off_i[1] -= e[1];
    }
  }
  if (contact_j == 1 || contact_j == 2) {
    double phi = A->orientation[j];
    double e[2] = {A->r_ts[j] * sin(phi), A->r_ts[j] * (-cos(phi))};
    if (contact_j == 1) {
      r_off_j[0] += e[0];
      r_off_j[1] += e[1];
    } else {
      r_off_j[0] -= e[0];
      r_off_j[1] -= e[1];
    }
  }
  const double *xi = A->position + 2 * i, *xj = A->position + 2 * j;
  double x_rel[2] = {xi[0] - xj[0], xi[1] - xj[1]};
  double r_off[2] = {r_off_i[0] - r_off_j[0], r_off_i[1] - r_off_j[1]};
  double grad[2];
  for (int k = 0; k < 2; ++k)
    grad[k] = (v_rel[k] - (a * (x_rel[k] + 2.0 * r_off[k]) + b_min * v_rel[k]) / d_min) / a;
  double si = (-A->mass[i]) * A->k_soc[i];
  double sj = (-A->mass[j]) * A->k_soc[j];
  double mi = magnitude(tau, A->tau_0[i]);
  double mj = magnitude(tau, A->tau_0[j]);
  for (int k = 0; k < 2; ++k) {
    fi[k] = 0.0 + (si * grad[k]) * mi;
    fj[k] = 0.0 - (sj * grad[k]) * mj;
  }
  truncate2(fi, ORC_F_SOC_MAX);
  truncate2(fj, ORC_F_SOC_MAX);
}

/* ------------------------------------------------------------------ */
/* interactions (core/interactions.py)                                 */
/* ------------------------------------------------------------------ */

/* Pair contributions of one interaction, before accumulation. Returned so
 * tests can inspect a single pair.  Returns 1 when the pair contributes
 * (h < SIGHT_SOC), else 0. */
int orc_pair_circular(const orc_agents *A, int64_t i, int64_t j, double *fi,
                      double *fj) {
  /* interaction_agent_agent_circular, core/interactions.py:51-70 */
  double h, n[2];
  distance_circles(A->position + 2 * i, A->radius[i], A->position + 2 * j,
                   A->radius[j], &h, n);
  if (!(h < ORC_SIGHT_SOC)) return 0;
  force_social_circular(A, i, j, fi, fj);
  if (h < 0.0) {
    double t[2] = {n[1], -n[0]}; /* rotate270 */
    const double *vi = A->velocity + 2 * i, *vj = A->velocity + 2 * j;
    double v[2] = {vi[0] - vj[0], vi[1] - vj[1]};
    double ci[2], cj[2];
    force_contact(h, n, v, t, A->mu[i], A->kappa[i], A->damping[i], ci);
    force_contact(h, n, v, t, A->mu[j], A->kappa[j], A->damping[j], cj);
    fi[0] += ci[0];
    fi[1] += ci[1];
    fj[0] -= cj[0];
    fj[1] -= cj[1];
  }
  return 1;
}

int orc_pair_three_circle(const orc_agents *A, int64_t i, int64_t j,
                          double *fi, double *fj, double *mi, double *mj) {
  /* interaction_agent_agent_three_circle, core/interactions.py:73-104 */
  const double *x_i[3] = {A->position + 2 * i, A->position_ls + 2 * i,
                          A->position_rs + 2 * i};
  const double *x_j[3] = {A->position + 2 * j, A->position_ls + 2 * j,
                          A->position_rs + 2 * j};
  double r_i[3] = {A->r_t[i], A->r_s[i], A->r_s[i]};
  double r_j[3] = {A->r_t[j], A->r_s[j], A->r_s[j]};
  double h, n[2], rmi[2], rmj[2];
  distance_three_circles(x_i, r_i, x_j, r_j, &h, n, rmi, rmj);
  if (!(h < ORC_SIGHT_SOC)) return 0;
  force_social_three_circle(A, i, j, fi, fj);
  if (h < 0.0) {
    double t[2] = {n[1], -n[0]};
    const double *vi = A->velocity + 2 * i, *vj = A->velocity + 2 * j;
    double v[2] = {vi[0] - vj[0], vi[1] - vj[1]};
    double ci[2], cj[2];
    force_contact(h, n, v, t, A->mu[i], A->kappa[i], A->damping[i], ci);
    force_contact(h, n, v, t, A->mu[j], A->kappa[j], A->damping[j], cj);
    fi[0] += ci[0];
    fi[1] += ci[1];
    fj[0] -= cj[0];
    fj[1] -= cj[1];
  }
  *mi = cross2(rmi, fi);
  *mj = cross2(rmj, fj);
  return 1;
}

/* Optional per-agent sum of |pair contribution| (and |torque contribution|),
 * the rounding scale of a force sum; set by orc_agent_agent_scale. */
static double *g_fscale = NULL, *g_tscale = NULL;

static inline void apply_pair(orc_agents *A, int64_t p, int64_t q) {
  int64_t i = p < q ? p : q, j = p < q ? q : p;
  if (A->model == 0) {
    double fi[2], fj[2];
    if (orc_pair_circular(A, i, j, fi, fj)) {
      if (g_fscale) {
        g_fscale[i] += hypot(fi[0], fi[1]);
        g_fscale[j] += hypot(fj[0], fj[1]);
      }
      A->force[2 * i] += fi[0];
      A->force[2 * i + 1] += fi[1];
      A->force[2 * j] += fj[0];
      A->force[2 * j + 1] += fj[1];
    }
  } else {
    double fi[2], fj[2], mi, mj;
    if (orc_pair_three_circle(A, i, j, fi, fj, &mi, &mj)) {
      if (g_fscale) {
        g_fscale[i] += hypot(fi[0], fi[1]);
        g_fscale[j] += hypot(fj[0], fj[1]);
        g_tscale[i] += fabs(mi);
        g_tscale[j] += fabs(mj);
      }
      A->force[2 * i] += fi[0];
      A->force[2 * i + 1] += fi[1];
      A->force[2 * j] += fj[0];
      A->force[2 * j + 1] += fj[1];
      A->torque[i] += mi;
      A->torque[j] += mj;
    }
  }
}

/* add_to_cells (cell_lists, called at core/interactions.py:192-193) under the
 * binning contract in this file's header.  Outputs are int64 like the
 * reference's i8 arrays.  Returns 0, or -1 when the grid would exceed
 * `cell_capacity` cells, or -2 for a non-finite position. */
int orc_add_to_cells(const double *pos, int64_t n, double cell_size,
                     double *origin, int64_t *shape, int64_t *cell_of,
                     int64_t *points_indices, int64_t *cells_count,
                     int64_t *cells_offset, int64_t cell_capacity) {
  if (n <= 0) {
    origin[0] = origin[1] = 0.0;
    shape[0] = shape[1] = 0;
    return 0;
  }
  double mn[2] = {pos[0], pos[1]}, mx[2] = {pos[0], pos[1]};
  for (int64_t i = 0; i < n; ++i) {
    for (int k = 0; k < 2; ++k) {
      double v = pos[2 * i + k];
      if (!isfinite(v)) return -2;
      if (v < mn[k]) mn[k] = v;
      if (v > mx[k]) mx[k] = v;
    }
  }
  origin[0] = mn[0];
  origin[1] = mn[1];
  int64_t nx = (int64_t)((mx[0] - mn[0]) / cell_size) + 1;
  int64_t ny = (int64_t)((mx[1] - mn[1]) / cell_size) + 1;
  shape[0] = nx;
  shape[1] = ny;
  if (nx <= 0 || ny <= 0 || nx > cell_capacity || ny > cell_capacity ||
      nx * ny > cell_capacity)
    return -1;
  int64_t C = nx * ny;
  memset(cells_count, 0, sizeof(int64_t) * C);
  for (int64_t i = 0; i < n; ++i) {
    int64_t lx = (int64_t)((pos[2 * i] - mn[0]) / cell_size);
    int64_t ly = (int64_t)((pos[2 * i + 1] - mn[1]) / cell_size);
    int64_t c = lx * ny + ly;
    cell_of[i] = c;
    cells_count[c] += 1;
  }
  int64_t acc = 0;
  for (int64_t c = 0; c < C; ++c) {
    cells_offset[c] = acc;
    acc += cells_count[c];
  }
  /* stable scatter: agents visited in index order */
  int64_t *cursor = (int64_t *)malloc(sizeof(int64_t) * C);
  memcpy(cursor, cells_offset, sizeof(int64_t) * C);
  for (int64_t i = 0; i < n; ++i) points_indices[cursor[cell_of[i]]++] = i;
  free(cursor);
  return 0;
}

/* iter_nearest_neighbors (cell_lists, called at core/interactions.py:152-154)
 * under the contract above.  Writes up to max_pairs (min,max) pairs to
 * `pairs` ([max_pairs][2]) and returns the total pair count. */
int64_t orc_iter_pairs(int64_t nx, int64_t ny, const int64_t *points_indices,
                       const int64_t *cells_count, const int64_t *cells_offset,
                       int64_t *pairs, int64_t max_pairs) {
  static const int64_t half[4][2] = {{0, 1}, {1, -1}, {1, 0}, {1, 1}};
  int64_t np_ = 0;
  for (int64_t c = 0; c < nx * ny; ++c) {
    int64_t cx = c / ny, cy = c % ny;
    const int64_t *seg = points_indices + cells_offset[c];
    int64_t cnt = cells_count[c];
    for (int64_t a = 0; a < cnt; ++a)
      for (int64_t b = a + 1; b < cnt; ++b) {
        if (np_ < max_pairs) {
          int64_t p = seg[a], q = seg[b];
          pairs[2 * np_] = p < q ? p : q;
          pairs[2 * np_ + 1] = p < q ? q : p;
        }
        ++np_;
      }
    for (int s = 0; s < 4; ++s) {
      int64_t ox = cx + half[s][0], oy = cy + half[s][1];
      if (ox < 0 || ox >= nx || oy < 0 || oy >= ny) continue;
      int64_t o = ox * ny + oy;
      const int64_t *seg2 = points_indices + cells_offset[o];
      for (int64_t a = 0; a < cnt; ++a)
        for (int64_t b = 0; b < cells_count[o]; ++b) {
          if (np_ < max_pairs) {
            int64_t p = seg[a], q = seg2[b];
            pairs[2 * np_] = p < q ? p : q;
            pairs[2 * np_ + 1] = p < q ? q : p;
          }
          ++np_;
        }
    }
  }
  return np_;
}

/* agent_agent_block_list, core/interactions.py:191-205 */
int orc_agent_agent_block_list(orc_agents *A, double cell_size) {
  int64_t n = A->n;
  if (n <= 0) return 0;
  if (A->model != 0 && A->model != 1) return -3; /* InvalidType */
  double origin[2];
  int64_t shape[2];
  int64_t *cell_of = (int64_t *)malloc(sizeof(int64_t) * n);
  int64_t *pidx = (int64_t *)malloc(sizeof(int64_t) * n);
  /* capacity sized from the bounding box */
  double mn[2] = {INFINITY, INFINITY}, mx[2] = {-INFINITY, -INFINITY};
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < 2; ++k) {
      double v = A->position[2 * i + k];
      if (v < mn[k]) mn[k] = v;
      if (v > mx[k]) mx[k] = v;
    }
  double cx = (mx[0] - mn[0]) / cell_size + 2.0, cy = (mx[1] - mn[1]) / cell_size + 2.0;
  if (!(cx * cy < 4e9)) {
    free(cell_of);
    free(pidx);
    return -1;
  }
  int64_t cap = (int64_t)(cx * cy) + 1;
  int64_t *cnt = (int64_t *)malloc(sizeof(int64_t) * cap);
  int64_t *off = (int64_t *)malloc(sizeof(int64_t) * cap);
  int rc = orc_add_to_cells(A->position, n, cell_size, origin, shape, cell_of,
                            pidx, cnt, off, cap);
  if (rc == 0) {
    static const int64_t half[4][2] = {{0, 1}, {1, -1}, {1, 0}, {1, 1}};
    int64_t nx = shape[0], ny = shape[1];
    for (int64_t c = 0; c < nx * ny; ++c) {
      int64_t ccx = c / ny, ccy = c % ny;
      const int64_t *seg = pidx + off[c];
      int64_t k = cnt[c];
      for (int64_t a = 0; a < k; ++a)
        for (int64_t b = a + 1; b < k; ++b) apply_pair(A, seg[a], seg[b]);
      for (int s = 0; s < 4; ++s) {
        int64_t ox = ccx + half[s][0], oy = ccy + half[s][1];
        if (ox < 0 || ox >= nx || oy < 0 || oy >= ny) continue;
        int64_t o = ox * ny + oy;
        const int64_t *seg2 = pidx + off[o];
        for (int64_t a = 0; a < k; ++a)
          for (int64_t b = 0; b < cnt[o]; ++b) apply_pair(A, seg[a], seg2[b]);
      }
    }
  }
  free(cell_of);
  free(pidx);
  free(cnt);
  free(off);
  return rc;
}

/* agent_agent_block_list that also returns, per agent, the sum of the norms
 * of its pair contributions (fscale) and of |torque contributions| (tscale):
 * the magnitude that bounds the rounding difference of two summation
 * orders. */
int orc_agent_agent_scale(orc_agents *A, double cell_size, double *fscale, double *tscale) {
  g_fscale = fscale;
  g_tscale = tscale;
  int rc = orc_agent_agent_block_list(A, cell_size);
  g_fscale = NULL;
  g_tscale = NULL;
  return rc;
}

/* agent_obstacle, core/interactions.py:107-141,169-186,208-214.
 * `segments` is obstacle_type_linear[W] = {p0: f8[2], p1: f8[2]}. */
int orc_agent_obstacle(orc_agents *A, const double *segments, int64_t W) {
#pragma omp parallel for schedule(static) if (g_par)
  for (int64_t i = 0; i < A->n; ++i) {
    for (int64_t w = 0; w < W; ++w) {
      const double *p0 = segments + 4 * w, *p1 = segments + 4 * w + 2;
      double h, n[2], rm[2] = {0.0, 0.0};
      if (A->model == 0) {
        orc_distance_circle_line(A->position + 2 * i, A->radius[i], p0, p1, &h, n);
      } else {
        const double *x[3] = {A->position + 2 * i, A->position_ls + 2 * i,
                              A->position_rs + 2 * i};
        double r[3] = {A->r_t[i], A->r_s[i], A->r_s[i]};
        distance_three_circle_line(x, r, p0, p1, &h, n, rm);
      }
      if (h < 0.0) {
        double t[2] = {n[1], -n[0]};
        double f[2];
        force_contact(h, n, A->velocity + 2 * i, t, A->mu[i], A->kappa[i],
                      A->damping[i], f);
        A->force[2 * i] += f[0];
        A->force[2 * i + 1] += f[1];
        if (A->model == 1) A->torque[i] += cross2(rm, f);
      }
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* adjusting, orientation, navigation                                   */
/* ------------------------------------------------------------------ */

/* force_adjust_agents, core/motion/adjusting.py:16-51,98-108 */
void orc_force_adjust_agents(orc_agents *A) {
#pragma omp parallel for schedule(static) if (g_par)
  for (int64_t i = 0; i < A->n; ++i) {
    double s = A->mass[i] / A->tau_adj[i];
    double v0 = A->target_velocity[i];
    for (int k = 0; k < 2; ++k)
      A->force[2 * i + k] += s * (v0 * A->target_direction[2 * i + k] - A->velocity[2 * i + k]);
  }
}

/* torque_adjust_agents, core/motion/adjusting.py:54-95,111-121 */
void orc_torque_adjust_agents(orc_agents *A) {
  for (int64_t i = 0; i < A->n; ++i) {
    double s = A->inertia_rot[i] / A->tau_rot[i];
    A->torque[i] += s * ((orc_wrap_to_pi(A->target_orientation[i] - A->orientation[i]) / M_PI) *
                             A->target_angular_velocity[i] -
                         A->angular_velocity[i]);
  }
}

/* orient_towards_target_direction, core/steering/orientation.py:17-21 */
void orc_orient_towards_target_direction(orc_agents *A) {
  for (int64_t i = 0; i < A->n; ++i)
    A->target_orientation[i] = atan2(A->target_direction[2 * i + 1], A->target_direction[2 * i]);
}

/* The same with the velocity as the direction (no reference counterpart:
 * a workload that aligns the body with its motion instead of its desired
 * direction). */
void orc_orient_towards_velocity(orc_agents *A) {
  for (int64_t i = 0; i < A->n; ++i)
    A->target_orientation[i] = atan2(A->velocity[2 * i + 1], A->velocity[2 * i]);
}

/* Navigation.update (simulation/logic.py:149-165) with meshgrid().indicer
 * (core/steering/quickest_path.py:41-44) and getdefault/is_inside
 * (core/steering/navigation.py:59-78).  `maps` holds n_targets direction
 * maps, each [rows][cols][2] = (u, v) interleaved.  Agents whose target is
 * outside [0, n_targets) are left unchanged (they match no target mask). */
void orc_navigation_nearest(orc_agents *A, const double *maps, int64_t n_targets,
                            int64_t rows, int64_t cols, double minx, double miny,
                            double step) {
#pragma omp parallel for schedule(static) if (g_par)
  for (int64_t i = 0; i < A->n; ++i) {
    int64_t t = A->target[i];
    if (t < 0 || t >= n_targets) continue;
    double qx = (A->position[2 * i] - minx) / step;
    double qy = (A->position[2 * i + 1] - miny) / step;
    /* astype(int64) truncates toward zero; an index is inside iff
     * 0 <= idx < shape, i.e. -1 < q < shape for the truncated value. */
    if (!(qx > -1.0 && qx < (double)cols && qy > -1.0 && qy < (double)rows)) continue;
    int64_t jj = (int64_t)qx, ii = (int64_t)qy;
    const double *m = maps + ((t * rows + ii) * cols + jj) * 2;
    A->target_direction[2 * i] = m[0];
    A->target_direction[2 * i + 1] = m[1];
  }
}

/* Bilinear-gradient navigation (extension; the reference reads the nearest
 * node of a direction map).  `pot` holds n_targets potentials [rows][cols];
 * node (r, c) sits at (minx + c*step, miny + r*step).  e0 = -grad/|grad| of
 * the bilinear interpolant of the agent's cell; agents outside the grid or
 * where the gradient vanishes keep their direction. */
void orc_navigation_bilinear(orc_agents *A, const double *pot, int64_t n_targets,
                             int64_t rows, int64_t cols, double minx, double miny,
                             double step) {
#pragma omp parallel for schedule(static) if (g_par)
  for (int64_t i = 0; i < A->n; ++i) {
    int64_t t = A->target[i];
    if (t < 0 || t >= n_targets) continue;
    double qx = (A->position[2 * i] - minx) / step;
    double qy = (A->position[2 * i + 1] - miny) / step;
    if (!(qx >= 0.0 && qx <= (double)(cols - 1) && qy >= 0.0 && qy <= (double)(rows - 1)))
      continue;
    int64_t c0 = (int64_t)qx, r0 = (int64_t)qy;
    if (c0 > cols - 2) c0 = cols - 2;
    if (r0 > rows - 2) r0 = rows - 2;
    if (c0 < 0 || r0 < 0) continue;
    double fx = qx - (double)c0, fy = qy - (double)r0;
    const double *P = pot + t * rows * cols;
    double p00 = P[r0 * cols + c0], p01 = P[r0 * cols + c0 + 1];
    double p10 = P[(r0 + 1) * cols + c0], p11 = P[(r0 + 1) * cols + c0 + 1];
    double gx = ((p01 - p00) * (1.0 - fy) + (p11 - p10) * fy) / step;
    double gy = ((p10 - p00) * (1.0 - fx) + (p11 - p01) * fx) / step;
    double g = hypot(gx, gy);
    if (!(g > 0.0)) continue;
    A->target_direction[2 * i] = -gx / g;
    A->target_direction[2 * i + 1] = -gy / g;
  }
}

/* ------------------------------------------------------------------ */
/* integrator (core/integrator.py)                                      */
/* ------------------------------------------------------------------ */

/* adaptive_timestep, core/integrator.py:29-97 */
double orc_adaptive_timestep(const orc_agents *A, double dt_min, double dt_max) {
  double v_max = 0.0;
#pragma omp parallel for schedule(static) reduction(max : v_max) if (g_par)
  for (int64_t i = 0; i < A->n; ++i) {
    double l = length2(A->velocity + 2 * i);
    if (l > v_max) v_max = l;
  }
  if (v_max == 0.0) return dt_max;
  double v0max = A->n > 0 ? A->target_velocity[0] : NAN;
  for (int64_t i = 1; i < A->n; ++i) {
    double v = A->target_velocity[i];
    if (isnan(v) || isnan(v0max)) v0max = NAN;
    else if (v > v0max) v0max = v;
  }
  double dx_max = 1.1 * v0max * dt_max;
  double dt = dx_max / v_max;
  if (dt > dt_max) return dt_max;
  if (dt < dt_min) return dt_min;
  return dt;
}

/* translational_verlet, core/integrator.py:164-176 */
void orc_translational_verlet(orc_agents *A, double dt) {
#pragma omp parallel for schedule(static) if (g_par)
  for (int64_t i = 0; i < A->n; ++i) {
    double m = A->mass[i];
    for (int k = 0; k < 2; ++k) {
      double a_old = A->force_prev[2 * i + k] / m;
      double a_new = A->force[2 * i + k] / m;
      A->force_prev[2 * i + k] = A->force[2 * i + k];
      A->velocity[2 * i + k] += (a_old + a_new) / 2.0 * dt;
      A->position[2 * i + k] += A->velocity[2 * i + k] * dt + a_new / 2.0 * (dt * dt);
    }
  }
}

/* rotational_verlet (core/integrator.py:179-193) and shoulders
 * (simulation/agents.py:473-486) */
void orc_rotational_verlet_shoulders(orc_agents *A, double dt) {
  for (int64_t i = 0; i < A->n; ++i) {
    double I = A->inertia_rot[i];
    double ao = A->torque_prev[i] / I;
    double an = A->torque[i] / I;
    A->torque_prev[i] = A->torque[i];
    A->angular_velocity[i] += (ao + an) / 2.0 * dt;
    A->orientation[i] += A->angular_velocity[i] * dt + an / 2.0 * (dt * dt);
    A->orientation[i] = orc_wrap_to_pi(A->orientation[i]);
  }
  for (int64_t i = 0; i < A->n; ++i) {
    double phi = A->orientation[i];
    double off[2] = {sin(phi) * A->r_ts[i], (-cos(phi)) * A->r_ts[i]};
    for (int k = 0; k < 2; ++k) {
      A->position_ls[2 * i + k] = A->position[2 * i + k] - off[k];
      A->position_rs[2 * i + k] = A->position[2 * i + k] + off[k];
    }
  }
}

/* velocity_verlet_integrator, core/integrator.py:209-256 */
double orc_velocity_verlet_integrator(orc_agents *A, double dt_min, double dt_max) {
  double dt = orc_adaptive_timestep(A, dt_min, dt_max);
  orc_translational_verlet(A, dt);
  if (A->model == 1) orc_rotational_verlet_shoulders(A, dt);
  return dt;
}

/* Reset.update, simulation/logic.py:59-64 */
void orc_reset(orc_agents *A) {
  /* memset is already bandwidth-bound */
  memset(A->force, 0, sizeof(double) * 2 * A->n);
  if (A->model == 1) memset(A->torque, 0, sizeof(double) * A->n);
}

/* ------------------------------------------------------------------ */
/* fluctuation (core/motion/fluctuation.py) with a counter-based stream */
/* ------------------------------------------------------------------ */

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). */
void orc_philox4x32(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

static inline double u01_from(uint32_t hi, uint32_t lo) {
  uint64_t x = (((uint64_t)hi << 32) | lo) >> 11;
  return ((double)x + 0.5) * 0x1.0p-53;
}

/* Draw for agent i, step s, stream `purpose` (0 = angle, 1 = force
 * magnitude, 2 = torque), attempt a.  Counter = (i, s_lo, s_hi, purpose |
 * a << 8); key = seed. */
static void orc_draw(uint64_t seed, int64_t i, uint64_t s, uint32_t purpose,
                     uint32_t attempt, double *u0, double *u1) {
  uint32_t ctr[4] = {(uint32_t)i, (uint32_t)s, (uint32_t)(s >> 32), purpose | (attempt << 8)};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t o[4];
  orc_philox4x32(ctr, key, o);
  *u0 = u01_from(o[0], o[1]);
  *u1 = u01_from(o[2], o[3]);
}

/* Standard normal truncated to [-3, 3] by Box-Muller with rejection. */
static double orc_truncnorm3(uint64_t seed, int64_t i, uint64_t s, uint32_t purpose) {
  double z = 0.0;
  for (uint32_t a = 0; a < 16; ++a) {
    double u0, u1;
    orc_draw(seed, i, s, purpose, a, &u0, &u1);
    z = sqrt(-2.0 * log(u0)) * cos(2.0 * M_PI * u1);
    if (fabs(z) <= 3.0) return z;
  }
  return copysign(3.0, z);
}

/* Per-agent noise of step s: force (2) and torque (1) increments.
 * force_fluctuation (core/motion/fluctuation.py:14-42):
 *   phi ~ U(0, 2pi), xi = scale * TN(0, 3), f = (mass * xi) * (cos, sin)
 * torque_fluctuation (core/motion/fluctuation.py:45-62):
 *   eta = inertia * (scale * TN(-3, 3)) */
void orc_fluctuation_noise(uint64_t seed, uint64_t s, int64_t i, double mass,
                           double scale_f, double inertia, double scale_t,
                           double *f_out, double *m_out) {
  double ua, ub;
  orc_draw(seed, i, s, 0, 0, &ua, &ub);
  double phi = 2.0 * M_PI * ua;
  double mag = scale_f * fabs(orc_truncnorm3(seed, i, s, 1));
  double mm = mass * mag;
  f_out[0] = mm * cos(phi);
  f_out[1] = mm * sin(phi);
  *m_out = inertia * (scale_t * orc_truncnorm3(seed, i, s, 2));
}

/* Fluctuation.update, simulation/logic.py:78-86 */
void orc_fluctuation(orc_agents *A, uint64_t seed, uint64_t s) {
#pragma omp parallel for schedule(static) if (g_par)
  for (int64_t i = 0; i < A->n; ++i) {
    double f[2], m;
    orc_fluctuation_noise(seed, s, i, A->mass[i], A->std_rand_force[i],
                          A->model == 1 ? A->inertia_rot[i] : 0.0,
                          A->model == 1 ? A->std_rand_torque[i] : 0.0, f, &m);
    A->force[2 * i] += f[0];
    A->force[2 * i + 1] += f[1];
    if (A->model == 1) A->torque[i] += m;
  }
}

/* ------------------------------------------------------------------ */
/* One step of the standard logic tree                                  */
/* ------------------------------------------------------------------ */

typedef struct orc_step_cfg {
  double cell_size, dt_min, dt_max;
  int32_t use_fluct, use_nav, nav_mode, use_adjust, use_agent_agent, use_obstacle, use_reset;
  int32_t orient_velocity; /* target_orientation = angle(velocity) (a workload's choice) */
  uint64_t seed;
  const double *segments;
  int64_t n_segments;
  const double *nav_maps; /* direction maps (mode 0) or potentials (mode 1) */
  int64_t nav_targets, nav_rows, nav_cols;
  double nav_minx, nav_miny, nav_step;
} orc_step_cfg;

/* Post-order of Reset << (Integrator << (Fluctuation, Adjusting << (Navigation,
 * Orientation), AgentAgentInteractions, AgentObstacleInteractions))
 * (examples/simulations.py:123-136, simulation/multiagent.py:51-55).
 * `step_index` feeds the fluctuation counter.  Returns dt (or NaN on a
 * binning error). */
double orc_step(orc_agents *A, const orc_step_cfg *cfg, uint64_t step_index) {
  if (cfg->use_fluct) orc_fluctuation(A, cfg->seed, step_index);
  if (cfg->use_nav) {
    if (cfg->nav_mode == 0)
      orc_navigation_nearest(A, cfg->nav_maps, cfg->nav_targets, cfg->nav_rows,
                             cfg->nav_cols, cfg->nav_minx, cfg->nav_miny, cfg->nav_step);
    else
      orc_navigation_bilinear(A, cfg->nav_maps, cfg->nav_targets, cfg->nav_rows,
                              cfg->nav_cols, cfg->nav_minx, cfg->nav_miny, cfg->nav_step);
  }
  if (A->model == 1 && cfg->use_adjust) {
    if (cfg->orient_velocity)
      orc_orient_towards_velocity(A);
    else
      orc_orient_towards_target_direction(A);
  }
  if (cfg->use_adjust) {
    orc_force_adjust_agents(A);
    if (A->model == 1) orc_torque_adjust_agents(A);
  }
  if (cfg->use_agent_agent) {
    if (orc_agent_agent_block_list(A, cfg->cell_size) != 0) return NAN;
  }
  if (cfg->use_obstacle) orc_agent_obstacle(A, cfg->segments, cfg->n_segments);
  double dt = orc_velocity_verlet_integrator(A, cfg->dt_min, cfg->dt_max);
  if (cfg->use_reset) orc_reset(A);
  return dt;
}

/* Runs `steps` steps; writes each dt to dts (may be NULL).  Returns 0 or a
 * negative error. */
int orc_run(orc_agents *A, const orc_step_cfg *cfg, uint64_t first_step, int64_t steps,
            double *dts) {
  for (int64_t s = 0; s < steps; ++s) {
    double dt = orc_step(A, cfg, first_step + (uint64_t)s);
    if (isnan(dt)) return -1;
    if (dts) dts[s] = dt;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* Multi-core variant of the agent-agent pass (host baseline only)      */
/* ------------------------------------------------------------------ */

/* Owner-computes over agents with OpenMP: each agent sums its own pair terms
 * (evaluated in canonical (min,max) order) over the 3x3 cell stencil.  Same
 * pair set and per-pair arithmetic as orc_agent_agent_block_list; only the
 * summation order differs.  Circular model only. */
int orc_agent_agent_owner_omp(orc_agents *A, double cell_size, int threads) {
  int64_t n = A->n;
  if (n <= 0) return 0;
  if (A->model != 0) return -3;
  double mn[2] = {INFINITY, INFINITY}, mx[2] = {-INFINITY, -INFINITY};
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < 2; ++k) {
      double v = A->position[2 * i + k];
      if (v < mn[k]) mn[k] = v;
      if (v > mx[k]) mx[k] = v;
    }
  double cx = (mx[0] - mn[0]) / cell_size + 2.0, cy = (mx[1] - mn[1]) / cell_size + 2.0;
  if (!(cx * cy < 4e9)) return -1;
  int64_t cap = (int64_t)(cx * cy) + 1;
  int64_t *cell_of = (int64_t *)malloc(sizeof(int64_t) * n);
  int64_t *pidx = (int64_t *)malloc(sizeof(int64_t) * n);
  int64_t *cnt = (int64_t *)malloc(sizeof(int64_t) * cap);
  int64_t *off = (int64_t *)malloc(sizeof(int64_t) * cap);
  double origin[2];
  int64_t shape[2];
  int rc = orc_add_to_cells(A->position, n, cell_size, origin, shape, cell_of, pidx, cnt, off, cap);
  if (rc == 0) {
    int64_t nx = shape[0], ny = shape[1];
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 256)
#endif
    for (int64_t i = 0; i < n; ++i) {
      int64_t c = cell_of[i], lx = c / ny, ly = c % ny;
      double acc[2] = {0.0, 0.0};
      for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy) {
          int64_t ox = lx + dx, oy = ly + dy;
          if (ox < 0 || ox >= nx || oy < 0 || oy >= ny) continue;
          int64_t o = ox * ny + oy;
          for (int64_t b = 0; b < cnt[o]; ++b) {
            int64_t j = pidx[off[o] + b];
            if (j == i) continue;
            double fi[2], fj[2];
            int64_t lo = i < j ? i : j, hi = i < j ? j : i;
            if (orc_pair_circular(A, lo, hi, fi, fj)) {
              const double *f = (i == lo) ? fi : fj;
              acc[0] += f[0];
              acc[1] += f[1];
            }
          }
        }
      A->force[2 * i] += acc[0];
      A->force[2 * i + 1] += acc[1];
    }
  }
  free(cell_of);
  free(pidx);
  free(cnt);
  free(off);
  return rc;
}

static double orc_step_omp_body(orc_agents *A, const orc_step_cfg *cfg, uint64_t step_index,
                                int threads);

/* One step like orc_step with the agent-agent pass and the other per-agent
 * passes spread over `threads` OpenMP threads (circular model). */
double orc_step_omp(orc_agents *A, const orc_step_cfg *cfg, uint64_t step_index, int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  g_par = 1;
  double dt = orc_step_omp_body(A, cfg, step_index, threads);
  g_par = 0;
  return dt;
}

static double orc_step_omp_body(orc_agents *A, const orc_step_cfg *cfg, uint64_t step_index,
                                int threads) {
  if (cfg->use_fluct) orc_fluctuation(A, cfg->seed, step_index);
  if (cfg->use_nav) {
    if (cfg->nav_mode == 0)
      orc_navigation_nearest(A, cfg->nav_maps, cfg->nav_targets, cfg->nav_rows,
                             cfg->nav_cols, cfg->nav_minx, cfg->nav_miny, cfg->nav_step);
    else
      orc_navigation_bilinear(A, cfg->nav_maps, cfg->nav_targets, cfg->nav_rows,
                              cfg->nav_cols, cfg->nav_minx, cfg->nav_miny, cfg->nav_step);
  }
  if (cfg->use_adjust) orc_force_adjust_agents(A);
  if (cfg->use_agent_agent) {
    if (orc_agent_agent_owner_omp(A, cfg->cell_size, threads) != 0) return NAN;
  }
  if (cfg->use_obstacle) orc_agent_obstacle(A, cfg->segments, cfg->n_segments);
  double dt = orc_velocity_verlet_integrator(A, cfg->dt_min, cfg->dt_max);
  if (cfg->use_reset) orc_reset(A);
  return dt;
}

int orc_run_omp(orc_agents *A, const orc_step_cfg *cfg, uint64_t first_step, int64_t steps,
                int threads, double *dts) {
  for (int64_t s = 0; s < steps; ++s) {
    double dt = orc_step_omp(A, cfg, first_step + (uint64_t)s, threads);
    if (isnan(dt)) return -1;
    if (dts) dts[s] = dt;
  }
  return 0;
}

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
