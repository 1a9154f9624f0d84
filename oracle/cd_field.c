/*
 * cd_field.c -- CPU oracle of the navigation field construction.  Test
 * infrastructure only: tests/, the smoke check and the CPU-baseline leg of
 * bench.py load it to check and time the device field builder; the product
 * never calls it.
 *
 * Restates, on a node grid of `rows` x `cols` (row = y, col = x, spacing
 * dx):
 *   - distance_map (crowddynamics/core/steering/quickest_path.py:54-115):
 *     scikit-fmm 0.0.9's distance(phi, dx) (requirements.txt pins it; the
 *     package is absent here), the signed distance to the zero contour of
 *     phi with masked nodes excluded.  scikit-fmm's published method: the
 *     nodes next to the contour are frozen at the distance given by linear
 *     interpolation of phi along each axis (d_k = dx phi_i / (phi_i -
 *     phi_j) at the nearest crossing, combined as 1 / sqrt(sum 1 / d_k^2)),
 *     then the fast marching method accepts nodes in order of distance,
 *     each from its accepted neighbours by the upwind discretisation of
 *     |grad u| = 1.  This restatement uses the first-order upwind update
 *     (scikit-fmm's default is second order where two upwind nodes are
 *     known), so that the device's parallel iteration, which converges to
 *     the fixed point of the same first-order update, is checked exactly.
 *   - direction_map (quickest_path.py:125-149): np.gradient of the map
 *     (central differences inside, one-sided at the edges, unit spacing),
 *     normalised; zero length gives NaN.
 *   - obstacle_handling (core/steering/obstacle_handling.py:15-74): within
 *     `radius` of the obstacles the target direction is blended with the
 *     direction away from them with weight p = strength^(x / radius), and
 *     every node is normalised.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* first-order upwind update from the smallest known neighbour values a
 * (along x) and b (along y); INFINITY marks an unknown side */
static double eik_update(double a, double b, double h) {
  if (a > b) {
    double t = a;
    a = b;
    b = t;
  }
  if (isinf(b) || b - a >= h) return a + h;
  double d = a - b;
  return (a + b + sqrt(2.0 * h * h - d * d)) / 2.0;
}

/* binary heap of (value, node) */
typedef struct {
  double *v;
  int64_t *id;
  int64_t n, cap;
} heap_t;

static void heap_push(heap_t *H, double v, int64_t id) {
  if (H->n == H->cap) {
    H->cap = H->cap ? 2 * H->cap : 1024;
    H->v = (double *)realloc(H->v, sizeof(double) * H->cap);
    H->id = (int64_t *)realloc(H->id, sizeof(int64_t) * H->cap);
  }
  int64_t k = H->n++;
  while (k > 0) {
    int64_t p = (k - 1) / 2;
    if (H->v[p] < v || (H->v[p] == v && H->id[p] < id)) break;
    H->v[k] = H->v[p];
    H->id[k] = H->id[p];
    k = p;
  }
  H->v[k] = v;
  H->id[k] = id;
}

static void heap_pop(heap_t *H, double *v, int64_t *id) {
  *v = H->v[0];
  *id = H->id[0];
  double lv = H->v[--H->n];
  int64_t lid = H->id[H->n];
  int64_t k = 0;
  for (;;) {
    int64_t c = 2 * k + 1;
    if (c >= H->n) break;
    if (c + 1 < H->n &&
        (H->v[c + 1] < H->v[c] || (H->v[c + 1] == H->v[c] && H->id[c + 1] < H->id[c])))
      ++c;
    if (lv < H->v[c] || (lv == H->v[c] && lid < H->id[c])) break;
    H->v[k] = H->v[c];
    H->id[k] = H->id[c];
    k = c;
  }
  H->v[k] = lv;
  H->id[k] = lid;
}

/* Signed distance (positive where phi > 0) to the zero contour of phi;
 * mask[k] != 0 excludes node k; excluded and unreachable nodes get NaN.
 * Returns 0, or -1 when no node borders the contour. */
int orc_distance_map(const double *phi, const uint8_t *mask, int64_t rows, int64_t cols,
                     double dx, double *out) {
  const int64_t N = rows * cols;
  uint8_t *state = (uint8_t *)calloc(N, 1); /* 0 far, 1 band, 2 accepted, 3 masked */
  double *u = (double *)malloc(sizeof(double) * N);
  heap_t H = {0, 0, 0, 0};
  int64_t frozen = 0;
  for (int64_t k = 0; k < N; ++k) {
    u[k] = INFINITY;
    if (mask && mask[k]) state[k] = 3;
  }
  /* nodes on the contour, then the nodes bordering it */
  for (int64_t k = 0; k < N; ++k)
    if (state[k] != 3 && phi[k] == 0.0) {
      u[k] = 0.0;
      state[k] = 2;
      ++frozen;
    }
  for (int64_t r = 0; r < rows; ++r)
    for (int64_t c = 0; c < cols; ++c) {
      const int64_t k = r * cols + c;
      if (state[k] != 0) continue;
      double ld[2] = {0.0, 0.0};
      int borders = 0;
      for (int dim = 0; dim < 2; ++dim)
        for (int s = -1; s <= 1; s += 2) {
          const int64_t rr = dim ? r + s : r, cc = dim ? c : c + s;
          if (rr < 0 || rr >= rows || cc < 0 || cc >= cols) continue;
          const int64_t j = rr * cols + cc;
          if (state[j] == 3 || !(phi[k] * phi[j] < 0.0)) continue;
          borders = 1;
          const double d = dx * phi[k] / (phi[k] - phi[j]);
          if (ld[dim] == 0.0 || ld[dim] > d) ld[dim] = d;
        }
      if (!borders) continue;
      double dsum = 0.0;
      for (int dim = 0; dim < 2; ++dim)
        if (ld[dim] > 0.0) dsum += 1.0 / ld[dim] / ld[dim];
      u[k] = sqrt(1.0 / dsum);
      state[k] = 2;
      ++frozen;
    }
  if (frozen == 0) {
    for (int64_t k = 0; k < N; ++k) out[k] = NAN;
    free(state);
    free(u);
    return -1;
  }
  /* the band around the frozen nodes */
  const int dr[4] = {0, 0, -1, 1}, dc[4] = {-1, 1, 0, 0};
#define NEIGHBOUR_MIN(r, c, dim, res)                                                   \
  do {                                                                                  \
    res = INFINITY;                                                                     \
    for (int s_ = -1; s_ <= 1; s_ += 2) {                                               \
      const int64_t rr_ = (dim) ? (r) + s_ : (r), cc_ = (dim) ? (c) : (c) + s_;         \
      if (rr_ < 0 || rr_ >= rows || cc_ < 0 || cc_ >= cols) continue;                   \
      const int64_t j_ = rr_ * cols + cc_;                                              \
      if (state[j_] == 2 && u[j_] < res) res = u[j_];                                   \
    }                                                                                   \
  } while (0)
  for (int64_t k = 0; k < N; ++k) {
    if (state[k] != 2) continue;
    const int64_t r = k / cols, c = k % cols;
    for (int q = 0; q < 4; ++q) {
      const int64_t rr = r + dr[q], cc = c + dc[q];
      if (rr < 0 || rr >= rows || cc < 0 || cc >= cols) continue;
      const int64_t j = rr * cols + cc;
      if (state[j] != 0) continue;
      double a, b;
      NEIGHBOUR_MIN(rr, cc, 0, a);
      NEIGHBOUR_MIN(rr, cc, 1, b);
      u[j] = eik_update(a, b, dx);
      state[j] = 1;
      heap_push(&H, u[j], j);
    }
  }
  while (H.n > 0) {
    double v;
    int64_t k;
    heap_pop(&H, &v, &k);
    if (state[k] == 2 || v != u[k]) continue; /* stale entry */
    state[k] = 2;
    const int64_t r = k / cols, c = k % cols;
    for (int q = 0; q < 4; ++q) {
      const int64_t rr = r + dr[q], cc = c + dc[q];
      if (rr < 0 || rr >= rows || cc < 0 || cc >= cols) continue;
      const int64_t j = rr * cols + cc;
      if (state[j] >= 2) continue;
      double a, b;
      NEIGHBOUR_MIN(rr, cc, 0, a);
      NEIGHBOUR_MIN(rr, cc, 1, b);
      const double w = eik_update(a, b, dx);
      if (state[j] == 0 || w < u[j]) {
        u[j] = w;
        state[j] = 1;
        heap_push(&H, w, j);
      }
    }
  }
#undef NEIGHBOUR_MIN
  for (int64_t k = 0; k < N; ++k) {
    if (state[k] != 2)
      out[k] = NAN;
    else
      out[k] = phi[k] < 0.0 ? -u[k] : u[k];
  }
  free(H.v);
  free(H.id);
  free(state);
  free(u);
  return 0;
}

/* direction_map: (ex, ey) = (d/dcol, d/drow) / |.| of np.gradient(dmap);
 * NaN inputs propagate, zero length gives NaN. */
void orc_direction_map(const double *dmap, int64_t rows, int64_t cols, double *ex, double *ey) {
  for (int64_t r = 0; r < rows; ++r)
    for (int64_t c = 0; c < cols; ++c) {
      const int64_t k = r * cols + c;
      double gu, gv; /* along rows (y), along cols (x) */
      if (rows < 2)
        gu = 0.0;
      else if (r == 0)
        gu = dmap[k + cols] - dmap[k];
      else if (r == rows - 1)
        gu = dmap[k] - dmap[k - cols];
      else
        gu = (dmap[k + cols] - dmap[k - cols]) / 2.0;
      if (cols < 2)
        gv = 0.0;
      else if (c == 0)
        gv = dmap[k + 1] - dmap[k];
      else if (c == cols - 1)
        gv = dmap[k] - dmap[k - 1];
      else
        gv = (dmap[k + 1] - dmap[k - 1]) / 2.0;
      double l = hypot(gu, gv);
      if (l == 0.0) l = NAN;
      ex[k] = gv / l;
      ey[k] = gu / l;
    }
}

/* obstacle_handling: blends (tx, ty) (direction to the targets) with the
 * direction away from the obstacles (-(ox, oy), (ox, oy) the direction map
 * of the obstacles' distance map dobs) where 0 < -dobs < radius, then
 * normalises every node; results in (tx, ty). */
void orc_obstacle_handling(const double *dobs, const double *ox, const double *oy, double *tx,
                           double *ty, int64_t n, double radius, double strength) {
  for (int64_t k = 0; k < n; ++k) {
    const double x = -dobs[k];
    if (0.0 < x && x < radius) {
      const double p = pow(strength, x / radius);
      tx[k] = -p * ox[k] + (1.0 - p) * tx[k];
      ty[k] = -p * oy[k] + (1.0 - p) * ty[k];
    }
  }
  for (int64_t k = 0; k < n; ++k) {
    const double l = hypot(tx[k], ty[k]);
    tx[k] = tx[k] / l;
    ty[k] = ty[k] / l;
  }
}
