// gvx_pba.hip -- DistanceVoxelMap: exact 3D EDT (PBA semantics), distance read-outs,
// mergeOccupied and the brute-force exactDistances3D.
//
// The EDT runs as three axis passes with no transposes of the map:
//   z pass  one thread per (x, y) column; the column's occupancy is held as a
//           register bitmask (all loads issued before any use), so each voxel is
//           read once and the nearest obstacle z (ties to the lower z) is written
//           as uint16.
//   y pass  one wave per 64 (x, z) lines (lanes = consecutive x).  The input is
//           staged through a 64 x 64 LDS tile; the proximate-point stack of PBA
//           phase 2 lives in a depth-major global scratch (stack[depth][line],
//           coalesced when lanes are at the same depth), then phase 3 colours the
//           line from the top.  Output: the packed DistanceVoxel (x, y_site,
//           z_site) written into the map.
//   x pass  one wave per 64 (y, z) lines (lanes = consecutive y), in place; the
//           LDS tile turns the per-lane x walk into contiguous row loads and row
//           stores.
// All bisector tests are exact int64 cross-multiplications.
#include "gvx_internal.hpp"

namespace gvx {

constexpr uint16_t ZNONE = 0xFFFF;

// ---------------------------------------------------------------- z pass
// FULL: dz == 32 * W, every load unconditional so the compiler keeps many in flight
template <int W, bool FULL>
__global__ void __launch_bounds__(256) k_pba_z(const uint32_t* __restrict__ map, uint16_t* __restrict__ zs,
                                               uint32_t dx, uint32_t dy, uint32_t dz, uint32_t uninit)
{
  const uint64_t col = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= (uint64_t)dx * dy) return;
  const uint64_t layer = (uint64_t)dx * dy;
  const uint32_t* __restrict__ p = map + col;
  uint32_t bits[W];
#pragma unroll
  for (int w = 0; w < W; ++w) {
    uint32_t v[32];
    if (FULL) {
#pragma unroll
      for (int b = 0; b < 32; ++b) v[b] = p[(uint64_t)(w * 32 + b) * layer];
    } else {
#pragma unroll
      for (int b = 0; b < 32; ++b) {
        const uint32_t z = (uint32_t)(w * 32 + b);
        v[b] = z < dz ? p[(uint64_t)z * layer] : uninit;
      }
    }
    uint32_t word = 0;
#pragma unroll
    for (int b = 0; b < 32; ++b) word |= (v[b] != uninit ? 1u : 0u) << b;
    bits[w] = word;
  }
  uint16_t* __restrict__ o = zs + col;
  uint32_t last = 0xFFFFu;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    // first set position in the words after w (65535 = none)
    uint32_t nxt = 0xFFFFu;
#pragma unroll
    for (int v = W - 1; v > w; --v) nxt = bits[v] ? (uint32_t)(v * 32 + __builtin_ctz(bits[v])) : nxt;
#pragma unroll
    for (int b = 0; b < 32; ++b) {
      const uint32_t z = (uint32_t)(w * 32 + b);
      if (FULL || z < dz) {
        if ((bits[w] >> b) & 1u) last = z;
        const uint32_t m = bits[w] & (0xFFFFFFFFu << b);
        const uint32_t hi = m ? (uint32_t)(w * 32 + __builtin_ctz(m)) : nxt;
        uint32_t pick;
        if (last == 0xFFFFu) pick = hi;
        else if (hi == 0xFFFFu) pick = last;
        else pick = (z - last <= hi - z) ? last : hi;
        o[(uint64_t)z * layer] = (uint16_t)pick;
      }
    }
  }
}

// columns longer than 1024: two sweeps through the uint16 buffer
__global__ void __launch_bounds__(256) k_pba_z_long(const uint32_t* __restrict__ map, uint16_t* __restrict__ zs,
                                                    uint32_t dx, uint32_t dy, uint32_t dz, uint32_t uninit)
{
  const uint64_t col = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= (uint64_t)dx * dy) return;
  const uint64_t layer = (uint64_t)dx * dy;
  uint32_t last = 0xFFFFu;
  for (uint32_t z = 0; z < dz; ++z) {
    if (map[z * layer + col] != uninit) last = z;
    zs[z * layer + col] = (uint16_t)last;
  }
  uint32_t next = 0xFFFFu;
  for (int64_t zz = (int64_t)dz - 1; zz >= 0; --zz) {
    const uint32_t z = (uint32_t)zz;
    if (map[z * layer + col] != uninit) next = z;
    const uint32_t lo = zs[z * layer + col];
    uint32_t pick;
    if (lo == 0xFFFFu) pick = next;
    else if (next == 0xFFFFu) pick = lo;
    else pick = (z - lo <= next - z) ? lo : next;
    zs[z * layer + col] = (uint16_t)pick;
  }
}

// exact bisector: for sites a < b on the line (position p, off-line squared cost c),
// Y(a,b) = N / D with N = (pb^2 - pa^2) + (cb - ca), D = 2 (pb - pa) > 0
__device__ __forceinline__ void bisector(int64_t pa, int64_t ca, int64_t pb, int64_t cb, int64_t& N, int64_t& D)
{
  N = (pb * pb - pa * pa) + (cb - ca);
  D = 2 * (pb - pa);
}

// ---------------------------------------------------------------- y and x passes
// One wave per block handles 64 lines.  Input values are staged through a 64 x 64
// LDS tile so every thread has 64 independent loads in flight (and, for the x
// pass, the global accesses are contiguous rows instead of one row per lane).
// The proximate-point stack of a line lives in stack[depth * nlines + line].
constexpr int TB = 64;

// y pass: lines (x, z), lanes = consecutive x; stack entry = y | zsite << 16;
// output: packed DistanceVoxel (x, y_site, z_site) written into the map
__global__ void __launch_bounds__(TB) k_pba_y(const uint16_t* __restrict__ zs, uint32_t* __restrict__ out,
                                              uint32_t* __restrict__ stack, uint32_t dx, uint32_t dy, uint32_t dz,
                                              uint32_t uninit)
{
  __shared__ uint16_t tile[TB][TB + 2];
  const uint32_t xbl = (dx + TB - 1) / TB;
  const uint32_t z = blockIdx.x / xbl, x0 = (blockIdx.x % xbl) * TB;
  const uint32_t lane = threadIdx.x, x = x0 + lane;
  const bool active = x < dx;
  const uint64_t nlines = (uint64_t)dx * dz;
  const uint64_t line = (uint64_t)z * dx + x;
  const uint64_t base = (uint64_t)z * dx * dy + x;
  const uint64_t cbase = (uint64_t)z * dx * dy + (active ? x : dx - 1);
  const int64_t zz = z;
  uint32_t n = 0, top = 0, sec = 0;
  for (uint32_t y0 = 0; y0 < dy; y0 += TB) {
    // clamped addresses: every load is unconditional so all of them are in flight at once
#pragma unroll 16
    for (int r = 0; r < TB; ++r) {
      const uint32_t yr = y0 + r < dy ? y0 + r : dy - 1;
      const uint16_t v = zs[cbase + (uint64_t)yr * dx];
      tile[r][lane] = (y0 + r < dy) ? v : ZNONE;  // inactive lanes read a clamped column they never use
    }
    __syncthreads();
    if (active) {
      const uint32_t lim = dy - y0 < (uint32_t)TB ? dy - y0 : (uint32_t)TB;
      for (uint32_t r = 0; r < lim; ++r) {
        const uint32_t s = tile[r][lane];
        if (s == ZNONE) continue;
        const uint32_t y = y0 + r;
        const int64_t c = ((int64_t)s - zz) * ((int64_t)s - zz);
        while (n >= 2) {
          const int64_t tp = top & 0xFFFF, tc = ((int64_t)(top >> 16) - zz) * ((int64_t)(top >> 16) - zz);
          const int64_t sp = sec & 0xFFFF, sc = ((int64_t)(sec >> 16) - zz) * ((int64_t)(sec >> 16) - zz);
          int64_t N1, D1, N2, D2;
          bisector(sp, sc, tp, tc, N1, D1);
          bisector(tp, tc, (int64_t)y, c, N2, D2);
          if (N1 * D2 < N2 * D1) break;
          --n;
          top = sec;
          if (n >= 2) sec = stack[(uint64_t)(n - 2) * nlines + line];
        }
        const uint32_t e = y | (s << 16);
        stack[(uint64_t)n * nlines + line] = e;
        ++n;
        sec = top;
        top = e;
      }
    }
    __syncthreads();
  }
  if (!active) return;
  if (n == 0) {
    for (uint32_t y = 0; y < dy; ++y) out[base + (uint64_t)y * dx] = uninit;
    return;
  }
  uint32_t k = n - 1;
  for (int64_t yy = (int64_t)dy - 1; yy >= 0; --yy) {
    while (k >= 1) {
      const int64_t tp = top & 0xFFFF, tc = ((int64_t)(top >> 16) - zz) * ((int64_t)(top >> 16) - zz);
      const int64_t sp = sec & 0xFFFF, sc = ((int64_t)(sec >> 16) - zz) * ((int64_t)(sec >> 16) - zz);
      int64_t N, D;
      bisector(sp, sc, tp, tc, N, D);
      if (yy * D <= N) {
        --k;
        top = sec;
        if (k >= 1) sec = stack[(uint64_t)(k - 1) * nlines + line];
      } else {
        break;
      }
    }
    out[base + (uint64_t)yy * dx] = x | ((top & 0xFFFFu) << 10) | ((top >> 16) << 20);
  }
}

// x pass (in place): lines (y, z), lanes = consecutive y; stack entry = the packed
// voxel value of the site (x | y << 10 | z << 20); tiles transpose rows through LDS
__global__ void __launch_bounds__(TB) k_pba_x(uint32_t* __restrict__ map, uint32_t* __restrict__ stack, uint32_t dx,
                                              uint32_t dy, uint32_t dz, uint32_t uninit)
{
  __shared__ uint32_t tile[TB][TB + 1];
  const uint32_t ybl = (dy + TB - 1) / TB;
  const uint32_t z = blockIdx.x / ybl, y0 = (blockIdx.x % ybl) * TB;
  const uint32_t lane = threadIdx.x, y = y0 + lane;
  const bool active = y < dy;
  const uint32_t rows = dy - y0 < (uint32_t)TB ? dy - y0 : (uint32_t)TB;
  const uint64_t nlines = (uint64_t)dy * dz;
  const uint64_t line = (uint64_t)z * dy + y;
  const uint64_t rowbase = ((uint64_t)z * dy + y0) * dx;
  const int64_t yq = y, zq = z;
  auto cost = [&](uint32_t e) -> int64_t {
    const int64_t sy = (e >> 10) & 0x3FF, sz = e >> 20;
    return (sy - yq) * (sy - yq) + (sz - zq) * (sz - zq);
  };
  const uint32_t ntiles = (dx + TB - 1) / TB;
  uint32_t n = 0, top = 0, sec = 0;
  for (uint32_t t = 0; t < ntiles; ++t) {
    const uint32_t x0 = t * TB, xr = x0 + lane;
    const uint32_t xc = xr < dx ? xr : dx - 1;
#pragma unroll 16
    for (uint32_t r = 0; r < (uint32_t)TB; ++r) {
      const uint32_t rc = r < rows ? r : rows - 1;
      const uint32_t v = map[rowbase + (uint64_t)rc * dx + xc];
      tile[r][lane] = (r < rows && xr < dx) ? v : uninit;
    }
    __syncthreads();
    if (active) {
      const uint32_t lim = dx - x0 < (uint32_t)TB ? dx - x0 : (uint32_t)TB;
      for (uint32_t lx = 0; lx < lim; ++lx) {
        const uint32_t e = tile[lane][lx];
        if (e == uninit) continue;
        const uint32_t x = x0 + lx;
        const int64_t c = cost(e);
        while (n >= 2) {
          int64_t N1, D1, N2, D2;
          bisector((int64_t)(sec & 0x3FF), cost(sec), (int64_t)(top & 0x3FF), cost(top), N1, D1);
          bisector((int64_t)(top & 0x3FF), cost(top), (int64_t)x, c, N2, D2);
          if (N1 * D2 < N2 * D1) break;
          --n;
          top = sec;
          if (n >= 2) sec = stack[(uint64_t)(n - 2) * nlines + line];
        }
        stack[(uint64_t)n * nlines + line] = e;
        ++n;
        sec = top;
        top = e;
      }
    }
    __syncthreads();
  }
  // colouring, tiles in descending x; results go through the tile to coalesced row stores
  uint32_t k = n ? n - 1 : 0;
  for (int64_t t = (int64_t)ntiles - 1; t >= 0; --t) {
    const uint32_t x0 = (uint32_t)t * TB, xr = x0 + lane;
    const uint32_t lim = dx - x0 < (uint32_t)TB ? dx - x0 : (uint32_t)TB;
    if (active) {
      for (int64_t lx = (int64_t)lim - 1; lx >= 0; --lx) {
        if (n == 0) { tile[lane][lx] = uninit; continue; }
        const int64_t xx = (int64_t)x0 + lx;
        while (k >= 1) {
          int64_t N, D;
          bisector((int64_t)(sec & 0x3FF), cost(sec), (int64_t)(top & 0x3FF), cost(top), N, D);
          if (xx * D <= N) {
            --k;
            top = sec;
            if (k >= 1) sec = stack[(uint64_t)(k - 1) * nlines + line];
          } else {
            break;
          }
        }
        tile[lane][lx] = top;
      }
    }
    __syncthreads();
    for (uint32_t r = 0; r < rows; ++r)
      if (xr < dx) map[rowbase + (uint64_t)r * dx + xr] = tile[r][lane];
    __syncthreads();
  }
}

// ---------------------------------------------------------------- read-outs
// DistanceVoxel::squaredObstacleDistance (DistanceVoxel.hpp:38-55)
__device__ __forceinline__ int32_t dist_sq(uint32_t raw, uint64_t id, uint32_t dx, uint32_t dy, uint32_t uninit)
{
  if (raw == uninit) return INT32_MAX;
  const int32_t x = (int32_t)(id % dx), y = (int32_t)((id / dx) % dy), z = (int32_t)(id / ((uint64_t)dx * dy));
  const int32_t ex = x - (int32_t)(raw & 0x3FF), ey = y - (int32_t)((raw >> 10) & 0x3FF), ez = z - (int32_t)(raw >> 20);
  return ex * ex + ey * ey + ez * ez;
}

__global__ void k_dist_sq(const uint32_t* __restrict__ map, const uint32_t* __restrict__ idx, uint64_t n, uint32_t dx,
                          uint32_t dy, uint32_t uninit, int32_t* __restrict__ out)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint64_t id = idx ? idx[i] : i;
    out[i] = dist_sq(map[id], id, dx, dy, uninit);
  }
}

__global__ void k_dist_metric(const uint32_t* __restrict__ map, uint64_t n, uint32_t dx, uint32_t dy, uint32_t uninit,
                              float side, float* __restrict__ out)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int32_t d = dist_sq(map[i], i, dx, dy, uninit);
    out[i] = d == INT32_MAX ? 3.402823466e+38f : sqrtf((float)d) * side;
  }
}

// mergeOccupied (DistanceVoxelMap.hpp:109-136 / DistanceVoxelMap.h:92-134)
__global__ void k_merge_occupied(uint32_t* __restrict__ dist, const int8_t* __restrict__ prob, uint64_t n, uint32_t dx,
                                 uint32_t dy, uint32_t ox, uint32_t oy, uint32_t oz, int thr)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if ((int)prob[i] > thr) {
      const uint32_t x = (uint32_t)(i % dx), y = (uint32_t)((i / dx) % dy), z = (uint32_t)(i / ((uint64_t)dx * dy));
      dist[i] = (x + ox) | ((y + oy) << 10) | ((z + oz) << 20);
    }
  }
}

// kernelExactDistances3D (VoxelMapOperations.hpp:896-990): obstacle list staged
// through LDS in tiles; strict < keeps the first minimum in list order.
constexpr int EX_BLOCK = 256;
__global__ void __launch_bounds__(EX_BLOCK) k_exact(uint32_t* __restrict__ map, uint64_t n, uint32_t dx, uint32_t dy,
                                                    uint32_t uninit, const float* __restrict__ obst, uint64_t no,
                                                    float side)
{
  __shared__ int32_t ob[EX_BLOCK][3];
  const uint64_t id = (uint64_t)blockIdx.x * EX_BLOCK + threadIdx.x;
  const bool active = id < n;
  int32_t x = 0, y = 0, z = 0, cur = 0, best = 0;
  int32_t bo[3] = {0, 0, 0};
  bool found = false;
  if (active) {
    x = (int32_t)(id % dx); y = (int32_t)((id / dx) % dy); z = (int32_t)(id / ((uint64_t)dx * dy));
    cur = dist_sq(map[id], id, dx, dy, uninit);
    best = cur;
  }
  for (uint64_t t = 0; t < no; t += EX_BLOCK) {
    __syncthreads();
    const uint64_t k = t + threadIdx.x;
    if (k < no) {
      ob[threadIdx.x][0] = quantize_signed(obst[3 * k], side);
      ob[threadIdx.x][1] = quantize_signed(obst[3 * k + 1], side);
      ob[threadIdx.x][2] = quantize_signed(obst[3 * k + 2], side);
    }
    __syncthreads();
    if (!active || cur == 0) continue;
    const int cnt = (int)((no - t) < (uint64_t)EX_BLOCK ? (no - t) : (uint64_t)EX_BLOCK);
    for (int s = 0; s < cnt; ++s) {
      const int32_t a = ob[s][0], b = ob[s][1], c = ob[s][2];
      if (a == x && b == y && c == z) continue;
      if (a == 1023 || b == 1023 || c == 1023) continue;
      const int32_t ex = x - a, ey = y - b, ez = z - c;
      const int32_t dd = ex * ex + ey * ey + ez * ez;
      if (dd < best) { best = dd; bo[0] = a; bo[1] = b; bo[2] = c; found = true; }
    }
  }
  if (active && found && best < cur) map[id] = (uint32_t)bo[0] | ((uint32_t)bo[1] << 10) | ((uint32_t)bo[2] << 20);
}

struct PbaScratch {
  void* zs = nullptr;
  void* stack = nullptr;
  uint64_t n = 0;
};
static PbaScratch& pba_scratch()
{
  static PbaScratch s;
  return s;
}

int pba_async(MapS* m)
{
  const uint32_t dx = m->dims[0], dy = m->dims[1], dz = m->dims[2];
  PbaScratch& sc = pba_scratch();
  if (sc.n < m->n) {
    GVX_CHECK(sync());
    dfree(sc.zs);
    dfree(sc.stack);
    sc.zs = sc.stack = nullptr;
    sc.n = 0;
    GVX_CHECK(dalloc(&sc.zs, m->n * 2));
    GVX_CHECK(dalloc(&sc.stack, m->n * 4));
    sc.n = m->n;
  }
  hipStream_t s = rt().stream;
  uint32_t* map = (uint32_t*)m->data;
  uint16_t* zs = (uint16_t*)sc.zs;
  uint32_t* st = (uint32_t*)sc.stack;
  const uint64_t cols = (uint64_t)dx * dy;
  {
    ProfScope ps("pba_z");
    dim3 g((unsigned)((cols + 255) / 256)), b(256);
    const uint32_t u = m->uninit;
#define ZL(W)                                                                                   \
  if (dz == 32 * W) hipLaunchKernelGGL((k_pba_z<W, true>), g, b, 0, s, map, zs, dx, dy, dz, u); \
  else hipLaunchKernelGGL((k_pba_z<W, false>), g, b, 0, s, map, zs, dx, dy, dz, u)
    if (dz <= 128) { ZL(4); }
    else if (dz <= 256) { ZL(8); }
    else if (dz <= 512) { ZL(16); }
    else if (dz <= 1024) { ZL(32); }
    else hipLaunchKernelGGL(k_pba_z_long, g, b, 0, s, map, zs, dx, dy, dz, u);
#undef ZL
    GVX_HIP(hipGetLastError());
  }
  {
    ProfScope ps("pba_y");
    const uint64_t blocks = (uint64_t)((dx + TB - 1) / TB) * dz;
    hipLaunchKernelGGL(k_pba_y, dim3((unsigned)blocks), dim3(TB), 0, s, zs, map, st, dx, dy, dz, m->uninit);
    GVX_HIP(hipGetLastError());
  }
  {
    ProfScope ps("pba_x");
    const uint64_t blocks = (uint64_t)((dy + TB - 1) / TB) * dz;
    hipLaunchKernelGGL(k_pba_x, dim3((unsigned)blocks), dim3(TB), 0, s, map, st, dx, dy, dz, m->uninit);
    GVX_HIP(hipGetLastError());
  }
  return GVX_OK;
}

}  // namespace gvx

using namespace gvx;

static int check_dist(gvx_map m)
{
  if (!m || !m->data) return fail(GVX_ERR_BAD_ARG, "null map handle");
  if (m->type != GVX_MT_DISTANCE_VOXELMAP) return fail(GVX_ERR_UNSUPPORTED, "not a DistanceVoxelMap");
  return GVX_OK;
}

extern "C" {

int gvx_dist_pba(gvx_map m)
{
  GVX_CHECK(check_dist(m));
  std::lock_guard<std::recursive_mutex> g(m->mu);
  std::lock_guard<std::recursive_mutex> g2(rt().mu);
  {
    ProfScope ps("pba");
    GVX_CHECK(pba_async(m));
  }
  return sync();
}

int gvx_dist_merge_occupied(gvx_map dist, gvx_map prob, const uint32_t offset[3], float threshold)
{
  GVX_CHECK(check_dist(dist));
  if (!prob || prob->type != GVX_MT_PROBAB_VOXELMAP) return fail(GVX_ERR_BAD_ARG, "mergeOccupied needs a ProbVoxelMap");
  if (memcmp(dist->dims, prob->dims, sizeof(dist->dims)) != 0)
    return fail(GVX_ERR_DIM_MISMATCH, "mergeOccupied: map dimensions differ");
  std::unique_lock<std::recursive_mutex> g1(dist->mu, std::defer_lock), g2(prob->mu, std::defer_lock);
  std::lock(g1, g2);
  hipLaunchKernelGGL(k_merge_occupied, dim3(grid_for(dist->n, 256, 8192)), dim3(256), 0, rt().stream,
                     (uint32_t*)dist->data, (const int8_t*)prob->data, dist->n, dist->dims[0], dist->dims[1],
                     offset ? offset[0] : 0u, offset ? offset[1] : 0u, offset ? offset[2] : 0u,
                     (int)float_to_probability(threshold));
  GVX_HIP(hipGetLastError());
  return sync();
}

int gvx_dist_exact(gvx_map m, const float* obstacles, uint64_t n, int mem)
{
  GVX_CHECK(check_dist(m));
  if (n && !obstacles) return fail(GVX_ERR_BAD_ARG, "null obstacles");
  std::lock_guard<std::recursive_mutex> g(m->mu);
  std::lock_guard<std::recursive_mutex> g2(rt().mu);
  const void* d = nullptr;
  GVX_CHECK(stage_input(obstacles, n * 12, mem, &d));
  hipLaunchKernelGGL(k_exact, dim3((unsigned)((m->n + EX_BLOCK - 1) / EX_BLOCK)), dim3(EX_BLOCK), 0, rt().stream,
                     (uint32_t*)m->data, m->n, m->dims[0], m->dims[1], m->uninit, (const float*)d, n, m->side);
  GVX_HIP(hipGetLastError());
  return sync();
}

int gvx_dist_squared(gvx_map m, const uint32_t* idx, uint64_t n, int32_t* out, int mem)
{
  GVX_CHECK(check_dist(m));
  if (!out) return fail(GVX_ERR_BAD_ARG, "null out");
  if (!idx) n = m->n;
  if (n == 0) return GVX_OK;
  std::lock_guard<std::recursive_mutex> g(m->mu);
  std::lock_guard<std::recursive_mutex> g2(rt().mu);
  // indices staged in front, results behind
  size_t ib = idx ? ((n * 4 + 255) & ~(size_t)255) : 0;
  void* st = nullptr;
  const uint32_t* didx = nullptr;
  int32_t* dout = out;
  if (idx && mem == GVX_MEM_DEVICE) didx = idx;
  size_t need = (idx && mem == GVX_MEM_HOST ? ib : 0) + (mem == GVX_MEM_HOST ? n * 4 : 0);
  if (need) GVX_CHECK(stage_reserve(need, &st));
  if (idx && mem == GVX_MEM_HOST) {
    GVX_HIP(hipMemcpyAsync(st, idx, n * 4, hipMemcpyHostToDevice, rt().stream));
    didx = (const uint32_t*)st;
  }
  if (mem == GVX_MEM_HOST) dout = (int32_t*)((char*)st + (idx ? ib : 0));
  if (idx) {
    // bounds-check indices on the way: invalid ids are rejected before launch
    std::vector<uint32_t> host;
    const uint32_t* hp = idx;
    if (mem == GVX_MEM_DEVICE) {
      host.resize(n);
      GVX_HIP(hipMemcpyAsync(host.data(), idx, n * 4, hipMemcpyDeviceToHost, rt().stream));
      GVX_CHECK(sync());
      hp = host.data();
    }
    for (uint64_t i = 0; i < n; ++i)
      if (hp[i] >= m->n) return fail(GVX_ERR_BAD_ARG, "voxel index out of range");
  }
  hipLaunchKernelGGL(k_dist_sq, dim3(grid_for(n, 256, 8192)), dim3(256), 0, rt().stream, (const uint32_t*)m->data,
                     didx, n, m->dims[0], m->dims[1], m->uninit, dout);
  GVX_HIP(hipGetLastError());
  if (mem == GVX_MEM_HOST) GVX_HIP(hipMemcpyAsync(out, dout, n * 4, hipMemcpyDeviceToHost, rt().stream));
  return sync();
}

int gvx_dist_metric(gvx_map m, float* out, int mem)
{
  GVX_CHECK(check_dist(m));
  if (!out) return fail(GVX_ERR_BAD_ARG, "null out");
  std::lock_guard<std::recursive_mutex> g(m->mu);
  std::lock_guard<std::recursive_mutex> g2(rt().mu);
  float* dout = out;
  if (mem == GVX_MEM_HOST) {
    void* st = nullptr;
    GVX_CHECK(stage_reserve(m->n * 4, &st));
    dout = (float*)st;
  }
  hipLaunchKernelGGL(k_dist_metric, dim3(grid_for(m->n, 256, 8192)), dim3(256), 0, rt().stream,
                     (const uint32_t*)m->data, m->n, m->dims[0], m->dims[1], m->uninit, m->side, dout);
  GVX_HIP(hipGetLastError());
  if (mem == GVX_MEM_HOST) GVX_HIP(hipMemcpyAsync(out, dout, m->n * 4, hipMemcpyDeviceToHost, rt().stream));
  return sync();
}

}  // extern "C"
