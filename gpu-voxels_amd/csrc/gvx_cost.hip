// gvx_cost.hip -- differentiable clearance cost on the EDT (no reference counterpart).
//
//   c(traj) = sum over waypoints, links, spheres of  w * max(0, r + eps - d(x))^2
//   x = T_l(q) * c_local,  T_l = base * A_0(q_0) ... A_{l-1}(q_{l-1})
//   d(x) = trilinear interpolation of the metric EDT sampled at voxel centres.
// The adjoint is hand-written: dc/dx = -2 w h grad d; dc/dq_j = z_j . sum_{l>j}
// (M_l - o_j x G_l) for revolute joints (z_j . sum G_l for prismatic), where
// G_l = sum dc/dx and M_l = sum x cross dc/dx over link l's spheres and (z_j, o_j)
// are the z axis and origin of frame T_j.  All reductions run in a fixed order,
// so results are deterministic.
#include "gvx_internal.hpp"

namespace gvx {

struct CostS {
  const float* grid = nullptr;
  uint32_t dims[3];
  float side, eps, w;
  uint32_t nl = 0, spl = 0;
  float* dh = nullptr;
  float* base = nullptr;
  float* spheres = nullptr;
  float* T = nullptr;        // per (traj, wp): nl x 12
  float* partial = nullptr;  // per (traj, wp)
  uint64_t cap = 0;          // (traj * wp) capacity
  std::recursive_mutex mu;
};

}  // namespace gvx

struct gvx_cost_s : gvx::CostS {};

namespace gvx {

__global__ void k_cost_fk(const float* __restrict__ dh, const float* __restrict__ base, const float* __restrict__ q,
                          uint64_t n, uint32_t nl, float* __restrict__ T)
{
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float t[12], a[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) t[k] = base[k];
  for (uint32_t l = 0; l < nl; ++l) {
    float* o = T + (p * nl + l) * 12;
#pragma unroll
    for (int k = 0; k < 12; ++k) o[k] = t[k];
    dh_matrix(dh + 5 * l, q[p * nl + l], a);
    affine_mul(t, a, t);
  }
}

// trilinear sample and gradient (samples at voxel centres, clamped at the border)
__device__ __forceinline__ float trilinear(const float* __restrict__ g, uint32_t dx, uint32_t dy, uint32_t dz, float side,
                                           float x, float y, float z, float& gx, float& gy, float& gz)
{
  const uint32_t d[3] = {dx, dy, dz};
  const float p[3] = {x, y, z};
  int i0[3];
  float t[3], s[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float u = p[a] / side - 0.5f;
    s[a] = 1.0f / side;
    const float hi = (float)(d[a] - 1);
    if (!(u >= 0.0f)) { u = 0.0f; s[a] = 0.0f; }
    if (u > hi) { u = hi; s[a] = 0.0f; }
    int i = (int)floorf(u);
    if (i > (int)d[a] - 2) i = (int)d[a] - 2;
    if (i < 0) i = 0;
    i0[a] = i;
    t[a] = u - (float)i;
  }
  const uint64_t sy = dx, sz = (uint64_t)dx * dy;
  const uint64_t b = (uint64_t)i0[2] * sz + (uint64_t)i0[1] * sy + (uint64_t)i0[0];
  const float c000 = g[b], c001 = g[b + 1], c010 = g[b + sy], c011 = g[b + sy + 1];
  const float c100 = g[b + sz], c101 = g[b + sz + 1], c110 = g[b + sz + sy], c111 = g[b + sz + sy + 1];
  const float tx = t[0], ty = t[1], tz = t[2];
  const float c00 = c000 * (1.f - tx) + c001 * tx, c01 = c010 * (1.f - tx) + c011 * tx;
  const float c10 = c100 * (1.f - tx) + c101 * tx, c11 = c110 * (1.f - tx) + c111 * tx;
  const float c0 = c00 * (1.f - ty) + c01 * ty, c1 = c10 * (1.f - ty) + c11 * ty;
  const float ddx0 = (c001 - c000) * (1.f - ty) + (c011 - c010) * ty;
  const float ddx1 = (c101 - c100) * (1.f - ty) + (c111 - c110) * ty;
  gx = (ddx0 * (1.f - tz) + ddx1 * tz) * s[0];
  gy = ((c01 - c00) * (1.f - tz) + (c11 - c10) * tz) * s[1];
  gz = (c1 - c0) * s[2];
  return c0 * (1.f - tz) + c1 * tz;
}

constexpr int CB = 512;  // threads per (trajectory, waypoint) block

// forward: one block per (trajectory, waypoint); fixed-order tree reduction
__global__ void __launch_bounds__(CB) k_cost_fwd(const float* __restrict__ grid, uint32_t dx, uint32_t dy, uint32_t dz,
                                                 float side, const float* __restrict__ T,
                                                 const float* __restrict__ spheres, uint32_t nl, uint32_t spl, float eps,
                                                 float w, float* __restrict__ partial)
{
  __shared__ float red[CB];
  const uint64_t tw = blockIdx.x;
  const uint32_t ns = nl * spl;
  float acc = 0.f;
  for (uint32_t k = threadIdx.x; k < ns; k += CB) {
    const uint32_t l = k / spl;
    const float* m = T + (tw * nl + l) * 12;
    const float* sp = spheres + 4 * (uint64_t)k;
    float x, y, z, gx, gy, gz;
    transform(m, sp[0], sp[1], sp[2], x, y, z);
    const float d = trilinear(grid, dx, dy, dz, side, x, y, z, gx, gy, gz);
    const float h = sp[3] + eps - d;
    if (h > 0.f) acc += w * h * h;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = CB / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[tw] = red[0];
}

__global__ void k_cost_sum(const float* __restrict__ partial, uint32_t ntraj, uint32_t nwp, float* __restrict__ out)
{
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntraj) return;
  float s = 0.f;
  for (uint32_t k = 0; k < nwp; ++k) s += partial[(uint64_t)t * nwp + k];
  out[t] = s;
}

// backward: one block per (trajectory, waypoint)
__global__ void __launch_bounds__(CB) k_cost_bwd(const float* __restrict__ grid, uint32_t dx, uint32_t dy, uint32_t dz,
                                                 float side, const float* __restrict__ T,
                                                 const float* __restrict__ spheres, const float* __restrict__ dh,
                                                 uint32_t nl, uint32_t spl, uint32_t nwp, float eps, float w,
                                                 const float* __restrict__ upstream, float* __restrict__ dq,
                                                 float* __restrict__ dxo)
{
  __shared__ float contrib[CB][6];
  __shared__ float link_sum[64][6];
  const uint64_t tw = blockIdx.x;
  const uint32_t ns = nl * spl;
  const float up = upstream ? upstream[tw / nwp] : 1.0f;
  for (uint32_t base = 0; base < ns; base += CB) {
    const uint32_t k = base + threadIdx.x;
    float c6[6] = {0, 0, 0, 0, 0, 0};
    if (k < ns) {
      const uint32_t l = k / spl;
      const float* m = T + (tw * nl + l) * 12;
      const float* sp = spheres + 4 * (uint64_t)k;
      float x, y, z, gx, gy, gz;
      transform(m, sp[0], sp[1], sp[2], x, y, z);
      const float d = trilinear(grid, dx, dy, dz, side, x, y, z, gx, gy, gz);
      const float h = sp[3] + eps - d;
      float ax = 0.f, ay = 0.f, az = 0.f;
      if (h > 0.f) {
        const float f = -2.0f * w * h * up;
        ax = f * gx; ay = f * gy; az = f * gz;
      }
      if (dxo) {
        float* o = dxo + (tw * ns + k) * 3;
        o[0] = ax; o[1] = ay; o[2] = az;
      }
      c6[0] = ax; c6[1] = ay; c6[2] = az;
      c6[3] = y * az - z * ay; c6[4] = z * ax - x * az; c6[5] = x * ay - y * ax;
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) contrib[threadIdx.x][c] = c6[c];
    __syncthreads();
    // per (link, component): sum this slice's spheres in index order
    for (uint32_t r = threadIdx.x; r < nl * 6; r += CB) {
      const uint32_t l = r / 6, c = r % 6;
      const uint32_t lo = l * spl, hi = lo + spl;
      float s = (base == 0) ? 0.f : link_sum[l][c];
      for (uint32_t kk = (lo > base ? lo : base); kk < hi && kk < base + CB && kk < ns; ++kk) s += contrib[kk - base][c];
      link_sum[l][c] = s;
    }
    __syncthreads();
  }
  // dq_j = z_j . sum_{l > j} (M_l - o_j x G_l)   (revolute), z_j . sum G_l (prismatic)
  for (uint32_t j = threadIdx.x; j < nl; j += CB) {
    const float* f = T + (tw * nl + j) * 12;
    const float zx = f[2], zy = f[6], zz = f[10], ox = f[3], oy = f[7], oz = f[11];
    float Gx = 0.f, Gy = 0.f, Gz = 0.f, Mx = 0.f, My = 0.f, Mz = 0.f;
    for (uint32_t l = j + 1; l < nl; ++l) {
      Gx += link_sum[l][0]; Gy += link_sum[l][1]; Gz += link_sum[l][2];
      Mx += link_sum[l][3]; My += link_sum[l][4]; Mz += link_sum[l][5];
    }
    float r;
    if (dh[5 * j + 4] != 0.0f) {
      r = zx * Gx + zy * Gy + zz * Gz;
    } else {
      const float cx = Mx - (oy * Gz - oz * Gy), cy = My - (oz * Gx - ox * Gz), cz = Mz - (ox * Gy - oy * Gx);
      r = zx * cx + zy * cy + zz * cz;
    }
    dq[tw * nl + j] = r;
  }
}

static int cost_reserve(CostS* c, uint64_t n)
{
  if (n <= c->cap) return GVX_OK;
  GVX_CHECK(sync());
  dfree(c->T);
  dfree(c->partial);
  c->T = c->partial = nullptr;
  c->cap = 0;
  GVX_CHECK(dalloc((void**)&c->T, n * c->nl * 48));
  GVX_CHECK(dalloc((void**)&c->partial, n * 4));
  c->cap = n;
  return GVX_OK;
}

}  // namespace gvx

using namespace gvx;

extern "C" {

int gvx_cost_create(const float* grid_dev, const uint32_t dims[3], float side, const float* dh, uint32_t n_links,
                    const float base[12], const float* spheres, uint32_t spl, float eps, float weight, gvx_cost* out)
{
  if (!out) return fail(GVX_ERR_BAD_ARG, "null out");
  *out = nullptr;
  GVX_CHECK(ensure_init());
  if (!grid_dev || !dims || !dh || !spheres) return fail(GVX_ERR_BAD_ARG, "null argument");
  if (dims[0] < 2 || dims[1] < 2 || dims[2] < 2 || !(side > 0.f)) return fail(GVX_ERR_BAD_ARG, "grid must be >= 2^3");
  if (n_links == 0 || n_links > 64 || spl == 0) return fail(GVX_ERR_BAD_ARG, "1..64 links and >= 1 sphere per link");
  gvx_cost c = new gvx_cost_s();
  c->grid = grid_dev;
  memcpy(c->dims, dims, 12);
  c->side = side;
  c->eps = eps;
  c->w = weight;
  c->nl = n_links;
  c->spl = spl;
  float ident[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
  int s;
  if ((s = dalloc((void**)&c->dh, n_links * 20)) != GVX_OK ||
      (s = dalloc((void**)&c->base, 48)) != GVX_OK ||
      (s = dalloc((void**)&c->spheres, (uint64_t)n_links * spl * 16)) != GVX_OK) {
    gvx_cost_destroy(c);
    return s;
  }
  hipStream_t st = rt().stream;
  if (hipMemcpyAsync(c->dh, dh, n_links * 20, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(c->base, base ? base : ident, 48, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(c->spheres, spheres, (uint64_t)n_links * spl * 16, hipMemcpyHostToDevice, st) != hipSuccess) {
    gvx_cost_destroy(c);
    return fail(GVX_ERR_HIP, "copy cost description");
  }
  if ((s = sync()) != GVX_OK) {
    gvx_cost_destroy(c);
    return s;
  }
  *out = c;
  return GVX_OK;
}

int gvx_cost_destroy(gvx_cost c)
{
  if (!c) return GVX_OK;
  (void)sync();
  void* bufs[] = {c->dh, c->base, c->spheres, c->T, c->partial};
  for (auto p : bufs) dfree(p);
  delete c;
  return GVX_OK;
}

int gvx_cost_forward(gvx_cost c, const float* q, uint32_t ntraj, uint32_t nwp, float* out)
{
  if (!c || !q || !out) return fail(GVX_ERR_BAD_ARG, "null argument");
  const uint64_t n = (uint64_t)ntraj * nwp;
  if (n == 0) return GVX_OK;
  std::lock_guard<std::recursive_mutex> g(c->mu);
  std::lock_guard<std::recursive_mutex> g2(rt().mu);
  GVX_CHECK(cost_reserve(c, n));
  hipStream_t s = rt().stream;
  ProfScope ps("cost_forward");
  hipLaunchKernelGGL(k_cost_fk, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, c->dh, c->base, q, n, c->nl, c->T);
  hipLaunchKernelGGL(k_cost_fwd, dim3((unsigned)n), dim3(CB), 0, s, c->grid, c->dims[0], c->dims[1], c->dims[2], c->side,
                     c->T, c->spheres, c->nl, c->spl, c->eps, c->w, c->partial);
  hipLaunchKernelGGL(k_cost_sum, dim3((ntraj + 63) / 64), dim3(64), 0, s, c->partial, ntraj, nwp, out);
  GVX_HIP(hipGetLastError());
  return sync();
}

int gvx_cost_backward(gvx_cost c, const float* q, uint32_t ntraj, uint32_t nwp, const float* upstream, float* dq,
                      float* dxo)
{
  if (!c || !q || !dq) return fail(GVX_ERR_BAD_ARG, "null argument");
  const uint64_t n = (uint64_t)ntraj * nwp;
  if (n == 0) return GVX_OK;
  std::lock_guard<std::recursive_mutex> g(c->mu);
  std::lock_guard<std::recursive_mutex> g2(rt().mu);
  GVX_CHECK(cost_reserve(c, n));
  hipStream_t s = rt().stream;
  ProfScope ps("cost_backward");
  hipLaunchKernelGGL(k_cost_fk, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, c->dh, c->base, q, n, c->nl, c->T);
  hipLaunchKernelGGL(k_cost_bwd, dim3((unsigned)n), dim3(CB), 0, s, c->grid, c->dims[0], c->dims[1], c->dims[2], c->side,
                     c->T, c->spheres, c->dh, c->nl, c->spl, nwp, c->eps, c->w, upstream, dq, dxo);
  GVX_HIP(hipGetLastError());
  return sync();
}

}  // extern "C"
