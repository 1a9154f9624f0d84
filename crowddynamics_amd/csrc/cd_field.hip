// Navigation field construction on the device: the distance map, direction
// map and obstacle blending of crowddynamics' quickest-path navigation
// (crowddynamics/core/steering/quickest_path.py:54-197,
// core/steering/obstacle_handling.py:15-74, cached per target by
// simulation/field.py:131-164).  The reference solves the Eikonal equation
// with scikit-fmm 0.0.9 (distance(phi, dx)) on the host; here it runs on the
// node grid in HBM.
//
// Distance map.  The nodes bordering the zero contour of phi are frozen at
// scikit-fmm's linear-interpolation distance (d_k = dx phi_i / (phi_i -
// phi_j) at the nearest crossing along axis k, combined as
// 1 / sqrt(sum 1 / d_k^2)); every other unmasked node converges to the
// fixed point of the first-order upwind update
//   u = a + dx                                  if b - a >= dx
//   u = (a + b + sqrt(2 dx^2 - (a - b)^2)) / 2   otherwise,
// a <= b the smaller neighbour values along the two axes.  That is the
// solution the fast marching method computes (the CPU oracle marches it in
// order of distance); on the GPU it is reached by relaxation: a workgroup
// takes a 32 x 32 tile with a one-node halo into LDS and sweeps it in
// red-black order until the tile stops changing, and the launches repeat
// until no tile changes.  Values only decrease, so every node's final value
// is the fixed point whatever the order of updates.
//
// Direction map: np.gradient (central differences inside, one-sided at the
// edges), normalised with the reference library's hypot; zero length gives
// NaN, and NaN (masked) inputs propagate.  Obstacle blending: within radius
// of the obstacles the target direction is mixed with the direction away
// from them with weight strength^(x / radius), then every node is
// normalised.  Missing directions in the buffer around obstacles take the
// direction of the nearest node that has one (quickest_path.py:152-166).

#include "cd_common.h"

namespace {

constexpr int kFT = 32;                // tile side in nodes
constexpr int kFH = kFT + 2;           // with the halo
constexpr int kFThreads = 256;         // 4 nodes per thread
constexpr int kFInner = 2 * kFT + 16;  // red-black sweeps per tile visit (bounded)

enum : uint8_t { F_FREE = 0, F_FROZEN = 2, F_MASKED = 3 };

__device__ __forceinline__ double eik_update(double a, double b, double h) {
  if (a > b) {
    const double t = a;
    a = b;
    b = t;
  }
  if (__builtin_isinf(b) || b - a >= h) return a + h;
  const double d = a - b;
  return (a + b + sqrt(2.0 * h * h - d * d)) / 2.0;
}

// Frozen nodes (on or next to the contour), free nodes at +inf, masked
// nodes excluded.
__global__ void __launch_bounds__(256) k_field_init(const double *phi, const uint8_t *mask,
                                                    int rows, int cols, double dx, double *u,
                                                    uint8_t *state) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= (int64_t)rows * cols) return;
  const int r = (int)(k / cols), c = (int)(k - (int64_t)r * cols);
  if (mask && mask[k]) {
    u[k] = __builtin_inf();
    state[k] = F_MASKED;
    return;
  }
  const double p = phi[k];
  if (p == 0.0) {
    u[k] = 0.0;
    state[k] = F_FROZEN;
    return;
  }
  double ld[2] = {0.0, 0.0};
  bool borders = false;
#pragma unroll
  for (int dim = 0; dim < 2; ++dim) {
#pragma unroll
    for (int s = -1; s <= 1; s += 2) {
      const int rr = dim ? r + s : r, cc = dim ? c : c + s;
      if (rr < 0 || rr >= rows || cc < 0 || cc >= cols) continue;
      const int64_t j = (int64_t)rr * cols + cc;
      if (mask && mask[j]) continue;
      const double q = phi[j];
      if (!(p * q < 0.0)) continue;
      borders = true;
      const double d = dx * p / (p - q);
      if (ld[dim] == 0.0 || ld[dim] > d) ld[dim] = d;
    }
  }
  if (!borders) {
    u[k] = __builtin_inf();
    state[k] = F_FREE;
    return;
  }
  double dsum = 0.0;
#pragma unroll
  for (int dim = 0; dim < 2; ++dim)
    if (ld[dim] > 0.0) dsum += 1.0 / ld[dim] / ld[dim];
  u[k] = sqrt(1.0 / dsum);
  state[k] = F_FROZEN;
}

// One visit of every tile: load the tile and its halo, relax the free nodes
// in red-black order until none changes (at most kFInner sweeps), store the
// tile, and flag the launch as changed if any node decreased.
__global__ void __launch_bounds__(kFThreads) k_field_relax(int rows, int cols, double dx,
                                                           double *u, const uint8_t *state,
                                                           int *changed) {
  __shared__ double su[kFH * kFH];
  __shared__ uint8_t ss[kFH * kFH];
  const int r0 = blockIdx.y * kFT, c0 = blockIdx.x * kFT;
  const int tid = threadIdx.x;
  for (int idx = tid; idx < kFH * kFH; idx += kFThreads) {
    const int rr = r0 - 1 + idx / kFH, cc = c0 - 1 + idx % kFH;
    if (rr >= 0 && rr < rows && cc >= 0 && cc < cols) {
      const int64_t g = (int64_t)rr * cols + cc;
      su[idx] = u[g];
      ss[idx] = state[g];
    } else {
      su[idx] = __builtin_inf();
      ss[idx] = F_MASKED;
    }
  }
  __syncthreads();
  bool any = false;
  for (int it = 0; it < kFInner; ++it) {
    bool ch = false;
#pragma unroll
    for (int colour = 0; colour < 2; ++colour) {
#pragma unroll
      for (int q = 0; q < (kFT * kFT) / kFThreads; ++q) {
        const int li = tid + q * kFThreads;
        const int lr = li / kFT, lc = li % kFT;
        if (((lr + lc) & 1) != colour) continue;
        const int idx = (lr + 1) * kFH + (lc + 1);
        if (ss[idx] != F_FREE) continue;
        const double a = fmin(su[idx - 1], su[idx + 1]);
        const double b = fmin(su[idx - kFH], su[idx + kFH]);
        const double w = eik_update(a, b, dx);
        if (w < su[idx]) {
          su[idx] = w;
          ch = true;
        }
      }
      __syncthreads();
    }
    ch = __syncthreads_or(ch);
    if (!ch) break;
    any = true;
  }
  for (int li = tid; li < kFT * kFT; li += kFThreads) {
    const int rr = r0 + li / kFT, cc = c0 + li % kFT;
    if (rr < rows && cc < cols) {
      const int idx = (li / kFT + 1) * kFH + (li % kFT + 1);
      if (ss[idx] == F_FREE) u[(int64_t)rr * cols + cc] = su[idx];
    }
  }
  if (any && tid == 0) atomicOr(changed, 1);
}

// Signed distance: negative where phi < 0; masked and unreached nodes NaN.
__global__ void __launch_bounds__(256) k_field_finish(const double *phi, int64_t n, double *u,
                                                      const uint8_t *state) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double v = u[k];
  if (state[k] == F_MASKED || __builtin_isinf(v))
    u[k] = __builtin_nan("");
  else
    u[k] = phi[k] < 0.0 ? -v : v;
}

__global__ void __launch_bounds__(256) k_direction_map(const double *d, int rows, int cols,
                                                       double *ex, double *ey) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= (int64_t)rows * cols) return;
  const int r = (int)(k / cols), c = (int)(k - (int64_t)r * cols);
  double gu, gv;  // along rows (y), along cols (x)
  if (rows < 2)
    gu = 0.0;
  else if (r == 0)
    gu = d[k + cols] - d[k];
  else if (r == rows - 1)
    gu = d[k] - d[k - cols];
  else
    gu = (d[k + cols] - d[k - cols]) / 2.0;
  if (cols < 2)
    gv = 0.0;
  else if (c == 0)
    gv = d[k + 1] - d[k];
  else if (c == cols - 1)
    gv = d[k] - d[k - 1];
  else
    gv = (d[k + 1] - d[k - 1]) / 2.0;
  double l = hypot_ref(gu, gv);
  if (l == 0.0) l = __builtin_nan("");
  ex[k] = gv / l;
  ey[k] = gu / l;
}

__global__ void __launch_bounds__(256) k_obstacle_blend(const double *dobs, const double *ox,
                                                        const double *oy, double *tx,
                                                        double *ty, int64_t n, double radius,
                                                        double strength) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= n) return;
  double x0 = tx[k], y0 = ty[k];
  const double x = -dobs[k];
  if (0.0 < x && x < radius) {
    const double p = pow(strength, x / radius);
    x0 = -p * ox[k] + (1.0 - p) * x0;
    y0 = -p * oy[k] + (1.0 - p) * y0;
  }
  const double l = hypot_ref(x0, y0);
  tx[k] = x0 / l;
  ty[k] = y0 / l;
}

// Missing directions (NaN) at nodes of `fill` take those of the nearest
// node with a direction within `reach` nodes (squared index distance, ties
// to the smaller row, then column); nodes with none in reach stay NaN.
// Reads (sx, sy), writes (dx_, dy_): a separate output, so the search sees
// only the original directions.
__global__ void __launch_bounds__(256) k_fill_nearest(const double *sx, const double *sy,
                                                      const uint8_t *fill, int rows, int cols,
                                                      int reach, double *dx_, double *dy_) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= (int64_t)rows * cols) return;
  const double x0 = sx[k], y0 = sy[k];
  dx_[k] = x0;
  dy_[k] = y0;
  if (!fill[k] || !(__builtin_isnan(x0) || __builtin_isnan(y0))) return;
  const int r = (int)(k / cols), c = (int)(k - (int64_t)r * cols);
  int best = 0x7fffffff;
  int64_t bj = -1;
  for (int rr = max(0, r - reach); rr <= min(rows - 1, r + reach); ++rr)
    for (int cc = max(0, c - reach); cc <= min(cols - 1, c + reach); ++cc) {
      const int64_t j = (int64_t)rr * cols + cc;
      if (__builtin_isnan(sx[j]) || __builtin_isnan(sy[j])) continue;
      const int d2 = (rr - r) * (rr - r) + (cc - c) * (cc - c);
      if (d2 < best) {  // row-major scan: the first of equal distances wins
        best = d2;
        bj = j;
      }
    }
  if (bj >= 0) {
    dx_[k] = sx[bj];
    dy_[k] = sy[bj];
  }
}

__global__ void k_clear_flag(int *flag) { *flag = 0; }

constexpr int64_t kMaxFieldNodes = (int64_t)1 << 30;

}  // namespace

extern "C" {

int64_t cd_field_work_bytes(int64_t rows, int64_t cols) {
  if (rows < 1 || cols < 1 || rows * cols > kMaxFieldNodes || rows > 0x7fffffff ||
      cols > 0x7fffffff)
    return CD_ERR_ARG;
  return align_up(rows * cols) + 256;
}

int cd_distance_map(const double *phi, const uint8_t *mask, int64_t rows, int64_t cols,
                    double dx, double *out, void *work, int64_t work_bytes,
                    cd_stream_t stream) {
  const int64_t need = cd_field_work_bytes(rows, cols);
  if (need < 0 || !phi || !out || !work || !(dx > 0.0)) return CD_ERR_ARG;
  if (work_bytes < need) return CD_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = rows * cols;
  uint8_t *state = (uint8_t *)work;
  int *changed = (int *)((char *)work + align_up(n));
  const unsigned nb = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_field_init, dim3(nb), dim3(256), 0, s, phi, mask, (int)rows, (int)cols,
                     dx, out, state);
  int rc = check_launch();
  if (rc) return rc;
  const dim3 tiles((unsigned)((cols + kFT - 1) / kFT), (unsigned)((rows + kFT - 1) / kFT));
  // every relaxation pass moves the front across at least one tile, so the
  // passes are bounded by the tiles a shortest path can visit
  const int64_t max_passes = 4 * ((int64_t)tiles.x * tiles.y) + 64;
  int h_changed = 1;
  int64_t pass = 0;
  for (; pass < max_passes && h_changed; ++pass) {
    hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(1), 0, s, changed);
    hipLaunchKernelGGL(k_field_relax, tiles, dim3(kFThreads), 0, s, (int)rows, (int)cols, dx,
                       out, (const uint8_t *)state, changed);
    if ((rc = check_launch())) return rc;
    if (hipMemcpyAsync(&h_changed, changed, sizeof(int), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return CD_ERR_HIP;
  }
  if (h_changed) return CD_ERR_HIP;  // not converged within the bound
  hipLaunchKernelGGL(k_field_finish, dim3(nb), dim3(256), 0, s, phi, n, out,
                     (const uint8_t *)state);
  return check_launch();
}

int cd_direction_map(const double *dmap, int64_t rows, int64_t cols, double *ex, double *ey,
                     cd_stream_t stream) {
  if (cd_field_work_bytes(rows, cols) < 0 || !dmap || !ex || !ey) return CD_ERR_ARG;
  const int64_t n = rows * cols;
  hipLaunchKernelGGL(k_direction_map, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, dmap, (int)rows, (int)cols, ex, ey);
  return check_launch();
}

int cd_obstacle_handling(const double *dobs, const double *ox, const double *oy, double *tx,
                         double *ty, int64_t n, double radius, double strength,
                         cd_stream_t stream) {
  if (n < 0 || n > kMaxFieldNodes || (n > 0 && (!dobs || !ox || !oy || !tx || !ty)) ||
      !(radius > 0.0))
    return CD_ERR_ARG;
  if (n == 0) return CD_OK;
  hipLaunchKernelGGL(k_obstacle_blend, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, dobs, ox, oy, tx, ty, n, radius, strength);
  return check_launch();
}

int cd_fill_missing(const double *sx, const double *sy, const uint8_t *fill, int64_t rows,
                    int64_t cols, int32_t reach, double *ex, double *ey, cd_stream_t stream) {
  if (cd_field_work_bytes(rows, cols) < 0 || !sx || !sy || !fill || !ex || !ey || reach < 0 ||
      reach > 1024 || ex == sx || ey == sy)
    return CD_ERR_ARG;
  const int64_t n = rows * cols;
  hipLaunchKernelGGL(k_fill_nearest, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, sx, sy, fill, (int)rows, (int)cols, (int)reach, ex, ey);
  return check_launch();
}

}  // extern "C"
