// Agent state nodes and the herding neighbours on the device, so that the
// full logic tree keeps its host synchronisation off the step:
//
//   InsideDomain   (crowddynamics/simulation/logic.py:343-358): active =
//                  position inside the domain polygon; data['inactive'] +=
//                  the number of agents whose state changed.
//   TargetReached  (logic.py:361-387): reached_by |= position inside the
//                  target polygon; data['target_i'] = sum(reached_by).
//   herding        (core/steering/collective_motion.py:69-154,240-279): the
//                  size_nearest_other nearest agents within sight whose line
//                  of sight crosses no obstacle (find_nearest_neighbors over
//                  the block list at cell size = sight), and the herding
//                  direction of the followers from the neighbours heading
//                  away (herding_interaction).
//
// Point in polygon is the crossings test (E. Haines, "Point in Polygon
// Strategies", Graphics Gems IV, 1994) that matplotlib's Path.contains_points
// (called by the reference) implements: a ray toward +x from the point
// toggles the state at each edge whose endpoints straddle the point's y
// (y_k >= y on one side only), where the crossing lies at or beyond the
// point; the side is decided by the sign of a cross product, without a
// division, and the polygon closes from its last vertex back to its first.

#include "cd_common.h"

namespace {

__device__ __forceinline__ bool crossings_inside(double tx, double ty, const double2 *v, int V) {
  if (!(__builtin_isfinite(tx) && __builtin_isfinite(ty)) || V < 1) return false;
  bool inside = false;
  double x0 = v[0].x, y0 = v[0].y;
  bool f0 = y0 >= ty;
  for (int k = 1; k <= V; ++k) {
    const double2 w = v[k < V ? k : 0];
    const bool f1 = w.y >= ty;
    if (f0 != f1) {
      if (((w.y - ty) * (x0 - w.x) >= (w.x - tx) * (y0 - w.y)) == f1) inside = !inside;
    }
    f0 = f1;
    x0 = w.x;
    y0 = w.y;
  }
  return inside;
}

__global__ void __launch_bounds__(256) k_inside_domain(const double2 *pos, int n,
                                                       const double2 *v, int V,
                                                       uint8_t *active,
                                                       unsigned long long *changed) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool ch = false;
  if (i < n) {
    const double2 p = pos[i];
    const uint8_t now = crossings_inside(p.x, p.y, v, V) ? 1 : 0;
    ch = (active[i] != 0) != (now != 0);
    active[i] = now;
  }
  const unsigned long long b = __ballot(ch);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(changed, (unsigned long long)__popcll(b));
}

__global__ void k_zero_u64(unsigned long long *x) { *x = 0ull; }

__global__ void __launch_bounds__(256) k_target_reached(const double2 *pos, int n,
                                                        const double2 *v, int V,
                                                        uint8_t *reached,
                                                        unsigned long long *count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool r = false;
  if (i < n) {
    const double2 p = pos[i];
    r = reached[i] != 0 || crossings_inside(p.x, p.y, v, V);
    reached[i] = r ? 1 : 0;
  }
  const unsigned long long b = __ballot(r);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(count, (unsigned long long)__popcll(b));
}

// line_intersect (core/geom2D.py:38-59): segments x0-x1 and y0-y1 cross
// (parallel segments do not).
__device__ __forceinline__ bool line_intersect(double2 x0, double2 x1, double2 y0, double2 y1) {
  const double ux = x1.x - x0.x, uy = x1.y - x0.y;
  const double vx = y1.x - y0.x, vy = y1.y - y0.y;
  const double bx = y0.x - x0.x, by = y0.y - x0.y;
  const double d = ux * vy - uy * vx;
  if (d == 0.0) return false;
  const double t0 = bx * vy - by * vx;
  const double t1 = bx * uy - by * ux;
  const double a = t0 / d, c = t1 / d;
  return 0.0 <= a && a <= 1.0 && 0.0 <= c && c <= 1.0;
}

constexpr int kMaxNearest = 16;

// The k nearest neighbours of agent i within sight (strictly closer than
// sight) whose line of sight crosses no segment, over the 3 x 3 cells of
// the block list (cells of side sight); ascending distance, ties to the
// smaller index, missing slots -1.
__global__ void __launch_bounds__(256) k_nearest(const double2 *pos, int n, double sight, int K,
                                                 const double4 *segs, int W, Work Wk,
                                                 const ScnState *ss, int64_t *nbr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double bd[kMaxNearest];
  int bj[kMaxNearest];
  for (int k = 0; k < K; ++k) {
    bd[k] = sight;
    bj[k] = -1;
  }
  const int c = Wk.cell_of[i];
  const int nx = ss->nx, ny = ss->ny;
  if (c >= 0 && ss->valid) {
    const double2 pi = pos[i];
    const int lx = c / ny, ly = c - lx * ny;
    for (int dx = -1; dx <= 1; ++dx) {
      const int ox = lx + dx;
      if (ox < 0 || ox >= nx) continue;
      for (int dy = -1; dy <= 1; ++dy) {
        const int oy = ly + dy;
        if (oy < 0 || oy >= ny) continue;
        const int cc = ox * ny + oy;
        for (int e = Wk.offset[cc]; e < Wk.offset[cc + 1]; ++e) {
          const int j = Wk.pidx[e];
          if (j == i) continue;
          const double2 pj = pos[j];
          const double l = hypot_ref(pi.x - pj.x, pi.y - pj.y);
          if (!(l < bd[K - 1]) && !(l == bd[K - 1] && bj[K - 1] >= 0 && j < bj[K - 1]))
            continue;
          bool blocked = false;
          for (int w = 0; w < W && !blocked; ++w) {
            const double4 s = segs[w];
            blocked = line_intersect(pi, pj, make_double2(s.x, s.y), make_double2(s.z, s.w));
          }
          if (blocked) continue;
          // insert (l, j) into the sorted list
          int k = K - 1;
          while (k > 0 && (bd[k - 1] > l || (bd[k - 1] == l && (bj[k - 1] < 0 || bj[k - 1] > j)))) {
            bd[k] = bd[k - 1];
            bj[k] = bj[k - 1];
            --k;
          }
          bd[k] = l;
          bj[k] = j;
        }
      }
    }
  }
  for (int k = 0; k < K; ++k) nbr[(int64_t)i * K + k] = bj[k];
}

__device__ __forceinline__ double2 normalize2(double2 v) {
  const double l = hypot_ref(v.x, v.y);
  return l != 0.0 ? make_double2(v.x / l, v.y / l) : v;
}

// herding_relationship (collective_motion.py:26-58), its first result: j
// is "heading away" from i unless cos(phi) < c_j < 1, c_j = -e_ij . v_j/|v_j|
// (both velocities non-zero), e_ij the unit vector from i to j.
__device__ __forceinline__ bool heading_away(double2 x1, double2 x2, double2 v1, double2 v2,
                                             double cos_phi) {
  if (hypot_ref(v1.x, v1.y) == 0.0 || hypot_ref(v2.x, v2.y) == 0.0) return false;
  const double2 e = normalize2(make_double2(x2.x - x1.x, x2.y - x1.y));
  const double2 n2 = normalize2(v2);
  const double cj = -(e.x * n2.x + e.y * n2.y);
  return !(cos_phi < cj && cj < 1.0);
}

// herding_interaction (collective_motion.py:110-154) of the followers over
// their neighbour lists (nearest first).
__global__ void __launch_bounds__(256) k_herding(const double2 *pos, const double2 *vel, int n,
                                                 const uint8_t *is_follower,
                                                 const int64_t *nbr, int K,
                                                 double weight_position, double cos_phi,
                                                 double2 *dir, uint8_t *has) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double2 d = make_double2(0.0, 0.0);
  uint8_t h = 0;
  if (is_follower[i]) {
    const double2 xi = pos[i], vi = vel[i];
    double mpx = 0.0, mpy = 0.0, mvx = 0.0, mvy = 0.0;
    int num = 0;
    for (int k = 0; k < K; ++k) {
      const int64_t j = nbr[(int64_t)i * K + k];
      if (j < 0) continue;
      const double2 xj = pos[j], vj = vel[j];
      if (heading_away(xi, xj, vi, vj, cos_phi)) {
        mpx += xj.x;
        mpy += xj.y;
        mvx += vj.x;
        mvy += vj.y;
        ++num;
      }
    }
    if (num > 0) {
      const double2 e0 = normalize2(make_double2(mpx / num - xi.x, mpy / num - xi.y));
      const double2 e1 = normalize2(make_double2(mvx, mvy));
      // weighted_average (core/vector2D.py:206-232), then normalize
      d = normalize2(make_double2(weight_position * e0.x + (1.0 - weight_position) * e1.x,
                                  weight_position * e0.y + (1.0 - weight_position) * e1.y));
      h = 1;
    }
  }
  dir[i] = d;
  has[i] = h;
}

}  // namespace

extern "C" {

int cd_inside_domain(const double *pos, int64_t n, const double *vertices, int64_t n_vertices,
                     uint8_t *active, uint64_t *inactive, cd_stream_t stream) {
  if (n < 0 || n > 0x7fffffff || n_vertices < 0 || n_vertices > (1 << 20) || !inactive ||
      (n > 0 && (!pos || !active)) || (n_vertices > 0 && !vertices))
    return CD_ERR_ARG;
  if (n == 0) return CD_OK;
  hipLaunchKernelGGL(k_inside_domain, dim3(grid_for(n, kBS)), dim3(kBS), 0, (hipStream_t)stream,
                     (const double2 *)pos, (int)n, (const double2 *)vertices, (int)n_vertices,
                     active, (unsigned long long *)inactive);
  return check_launch();
}

int cd_target_reached(const double *pos, int64_t n, const double *vertices, int64_t n_vertices,
                      uint8_t *reached_by, uint64_t *count, cd_stream_t stream) {
  if (n < 0 || n > 0x7fffffff || n_vertices < 0 || n_vertices > (1 << 20) || !count ||
      (n > 0 && (!pos || !reached_by)) || (n_vertices > 0 && !vertices))
    return CD_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, (unsigned long long *)count);
  if (n > 0)
    hipLaunchKernelGGL(k_target_reached, dim3(grid_for(n, kBS)), dim3(kBS), 0, s,
                       (const double2 *)pos, (int)n, (const double2 *)vertices, (int)n_vertices,
                       reached_by, (unsigned long long *)count);
  return check_launch();
}

int cd_nearest_neighbors(const cd_agents_t *a, double sight, int32_t k, const double *segments,
                         int64_t n_segments, int64_t *neighbors, const cd_workspace_t *w,
                         cd_stream_t stream) {
  int rc = check_agents(a);
  if (rc) return rc;
  if (!(sight > 0.0) || k < 1 || k > kMaxNearest || n_segments < 0 || n_segments > 0x7fffffff ||
      (n_segments > 0 && !segments) || (a->n > 0 && !neighbors))
    return CD_ERR_ARG;
  if ((rc = check_ws(a, w))) return rc;
  if (a->n == 0) return CD_OK;
  // the block list at cell size = sight (collective_motion.py:250-257)
  if ((rc = cd_add_to_cells(a, sight, w, stream))) return rc;
  const Work W = carve(w);
  const Layout L = make_layout(w);
  hipLaunchKernelGGL(k_nearest, dim3(grid_for(a->n, kBS)), dim3(kBS), 0, (hipStream_t)stream,
                     (const double2 *)a->position, (int)a->n, sight, (int)k,
                     (const double4 *)segments, (int)n_segments, W,
                     (const ScnState *)((char *)w->ptr + L.scn_state), neighbors);
  return check_launch();
}

int cd_herding_interaction(const double *pos, const double *vel, int64_t n,
                           const uint8_t *is_follower, const int64_t *neighbors, int32_t k,
                           double weight_position, double cos_phi, double *direction,
                           uint8_t *has_direction, cd_stream_t stream) {
  if (n < 0 || n > 0x7fffffff || k < 1 || k > kMaxNearest ||
      (n > 0 && (!pos || !vel || !is_follower || !neighbors || !direction || !has_direction)))
    return CD_ERR_ARG;
  if (n == 0) return CD_OK;
  hipLaunchKernelGGL(k_herding, dim3(grid_for(n, kBS)), dim3(kBS), 0, (hipStream_t)stream,
                     (const double2 *)pos, (const double2 *)vel, (int)n, is_follower, neighbors,
                     (int)k, weight_position, cos_phi, (double2 *)direction, has_direction);
  return check_launch();
}

}  // extern "C"
