// gvx_robot.hip -- DH robot on the GPU: forward kinematics, link-cloud transform and the
// fused "transform + insert" of KinematicChain::setConfiguration + insertRobotIntoMap
// (robot/dh_robot/KinematicChain.cu:93-126, helpers/MetaPointCloud.cu:636-661,
// GpuVoxels.cpp:499-517): one FK thread and one insert launch instead of one transform
// launch and one device sync per link.
#include "gvx_internal.hpp"

namespace gvx {

int list_insert_device(void* list, const float* pts_dev, uint64_t n, const uint32_t* ends_dev, uint32_t ncl,
                       const uint32_t* meanings_dev, uint32_t meaning);

// link l posed by base * A_0(q_0) ... A_{l-1}(q_{l-1}) into T (nl x 12, any memory):
// the DH matrices in parallel into sA (LDS, nl x 12), then one thread chains them
__device__ __forceinline__ void block_fk(const float* __restrict__ dh, const float* __restrict__ base,
                                         const float* __restrict__ q, uint32_t nl, float* sA, float* T)
{
  for (uint32_t l = threadIdx.x; l < nl; l += blockDim.x) dh_matrix(dh + 5 * l, q[l], sA + 12 * l);
  __syncthreads();
  if (threadIdx.x == 0) {
    float t[12];
    for (int k = 0; k < 12; ++k) t[k] = base[k];
    for (uint32_t l = 0; l < nl; ++l) {
      for (int k = 0; k < 12; ++k) T[12 * l + k] = t[k];
      affine_mul(t, sA + 12 * l, t);
    }
  }
  __syncthreads();
}

// dynamic LDS: nl x 12 floats
__global__ void __launch_bounds__(256) k_robot_fk(const float* __restrict__ dh, const float* __restrict__ base,
                                                  const float* __restrict__ q, uint32_t nl, float* __restrict__ T)
{
  extern __shared__ float lds[];
  block_fk(dh, base, q, nl, lds, T);
}

__device__ __forceinline__ uint32_t find_link(const uint32_t* __restrict__ ends, uint32_t nl, uint64_t i)
{
  uint32_t lo = 0, hi = nl - 1;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if ((uint64_t)ends[mid] > i) hi = mid; else lo = mid + 1;
  }
  return lo;
}

// kernelTransformCloud (helpers/kernels/MetaPointCloudOperations.cu:78) for all links at once
__global__ void __launch_bounds__(256) k_robot_transform(const float* __restrict__ pts, const uint32_t* __restrict__ ends,
                                                         uint32_t nl, uint64_t n, const float* __restrict__ T,
                                                         float* __restrict__ out)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t l = find_link(ends, nl, i);
    float x, y, z;
    transform(T + 12 * l, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], x, y, z);
    out[3 * i] = x;
    out[3 * i + 1] = y;
    out[3 * i + 2] = z;
  }
}

// FK, transform and BitVector insert in one launch (insertRobotIntoMap of a BitVector
// map: the link meanings are set with atomicOr, which commutes, so the per-link
// order of the reference does not matter).  Every block poses the chain in LDS (dynamic
// LDS: 2 x nl x 12 floats + 2 x nl words), then inserts its points; lanes of a wave with
// the same (voxel, meaning) set the bit once.
__global__ void __launch_bounds__(256) k_robot_insert_bit(uint32_t* __restrict__ map, uint32_t dx, uint32_t dy,
                                                          uint32_t dz, float side, const float* __restrict__ pts,
                                                          const uint32_t* __restrict__ ends,
                                                          const uint32_t* __restrict__ meanings, uint32_t nl,
                                                          uint64_t n, const float* __restrict__ dh,
                                                          const float* __restrict__ base, const float* __restrict__ q,
                                                          int* __restrict__ outside)
{
  extern __shared__ float lds[];
  float* sA = lds;
  float* sT = lds + 12 * nl;
  uint32_t* sE = reinterpret_cast<uint32_t*>(lds + 24 * nl);
  uint32_t* sM = sE + nl;
  for (uint32_t l = threadIdx.x; l < nl; l += blockDim.x) {
    sE[l] = ends[l];
    sM[l] = meanings[l];
  }
  block_fk(dh, base, q, nl, sA, sT);
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    uint32_t lo = 0, hi = nl - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if ((uint64_t)sE[mid] > i) hi = mid; else lo = mid + 1;
    }
    float wx, wy, wz;
    transform(sT + 12 * lo, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], wx, wy, wz);
    const uint32_t x = quantize(wx, side), y = quantize(wy, side), z = quantize(wz, side);
    const bool in = x < dx && y < dy && z < dz;
    const uint32_t m = sM[lo];
    const uint64_t idx = in ? ((uint64_t)z * dy + y) * dx + x : 0ull;
    if (wave_run_head(in, idx, m)) atomicOr(map + idx * 8 + (m >> 5), 1u << (m & 31));
    if (wave_ballot(!in) && outside) *outside = 1;
  }
}

int insert_points_async(MapS* m, const float* pts_dev, uint64_t n, const uint32_t* ends_dev, uint32_t ncl,
                        const uint32_t* meanings_dev, uint32_t meaning, int* outside_dev);
int selfcoll_insert_async(MapS* m, const float* pts_dev, const uint32_t* sizes, uint32_t ncl,
                          const uint32_t* meanings, const uint8_t* masks, uint32_t* acc_dev);
int selfcoll_finish(const uint32_t* acc_dev, uint8_t colliding[32], int* self_collision, int* points_outside);

// scratch layout in the staging buffer: [T (nl*12 f) | dh (nl*5 f) | base (12 f) | q (nl f) |
//                                        ends (nl u32) | meanings (nl u32) | points in (n*3 f) | points out (n*3 f)]
struct RobotStage {
  float *T, *dh, *base, *q, *pin, *pout;
  uint32_t *ends, *meanings;
};

static int robot_stage(const float* pts, const uint32_t* sizes, uint32_t nl, const float* dh, const float* base,
                       const float* q, const uint32_t* meanings, uint32_t meaning, int mem, uint64_t* n_out,
                       RobotStage& S)
{
  if (nl > 1024) return fail(GVX_ERR_UNSUPPORTED, "more than 1024 links");  // the chain is posed in LDS
  std::vector<uint32_t> ends(nl), mn(nl);
  uint64_t n = 0;
  for (uint32_t l = 0; l < nl; ++l) {
    n += sizes[l];
    if (n > 0xFFFFFFFFull) return fail(GVX_ERR_UNSUPPORTED, "too many link points");
    ends[l] = (uint32_t)n;
    mn[l] = meanings ? meanings[l] : meaning;
    if (mn[l] > 255) return fail(GVX_ERR_BAD_ARG, "meaning > 255");
  }
  if (n && !pts) return fail(GVX_ERR_BAD_ARG, "null link points");
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  size_t oT = 0, oD = oT + al(nl * 48), oB = oD + al(nl * 20), oQ = oB + al(48), oE = oQ + al(nl * 4),
         oM = oE + al(nl * 4), oP = oM + al(nl * 4), oO = oP + al(n * 12), tot = oO + al(n * 12);
  void* st = nullptr;
  GVX_CHECK(stage_reserve(tot, &st));
  char* b = (char*)st;
  S.T = (float*)(b + oT); S.dh = (float*)(b + oD); S.base = (float*)(b + oB); S.q = (float*)(b + oQ);
  S.ends = (uint32_t*)(b + oE); S.meanings = (uint32_t*)(b + oM); S.pin = (float*)(b + oP); S.pout = (float*)(b + oO);
  hipStream_t s = rt().stream;
  float ident[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
  // small descriptors always come from the host, in one copy; points and joint values follow `mem`
  std::vector<uint8_t> hd(oP - oD, 0);
  memcpy(hd.data() + (oD - oD), dh, nl * 20);
  memcpy(hd.data() + (oB - oD), base ? base : ident, 48);
  memcpy(hd.data() + (oE - oD), ends.data(), nl * 4);
  memcpy(hd.data() + (oM - oD), mn.data(), nl * 4);
  if (mem == GVX_MEM_HOST) memcpy(hd.data() + (oQ - oD), q, nl * 4);
  GVX_HIP(hipMemcpyAsync(S.dh, hd.data(), hd.size(), hipMemcpyHostToDevice, s));
  if (mem != GVX_MEM_HOST) GVX_CHECK(copy_in(S.q, q, nl * 4, mem));
  if (mem == GVX_MEM_HOST) {
    if (n) GVX_HIP(hipMemcpyAsync(S.pin, pts, n * 12, hipMemcpyHostToDevice, s));
  } else {
    S.pin = const_cast<float*>(pts);
  }
  *n_out = n;
  return GVX_OK;
}

static int robot_pose(const RobotStage& S, uint32_t nl, uint64_t n)
{
  hipStream_t s = rt().stream;
  hipLaunchKernelGGL(k_robot_fk, dim3(1), dim3(64), nl * 48, s, S.dh, S.base, S.q, nl, S.T);
  if (n)
    hipLaunchKernelGGL(k_robot_transform, dim3(grid_for(n, 256, 1u << 14)), dim3(256), 0, s, S.pin, S.ends, nl, n,
                       S.T, S.pout);
  GVX_HIP(hipGetLastError());
  return GVX_OK;
}

}  // namespace gvx

using namespace gvx;

extern "C" {

int gvx_dh_matrix(const float dh[5], float value, float out[12])
{
  if (!dh || !out) return fail(GVX_ERR_BAD_ARG, "null argument");
  dh_matrix(dh, value, out);
  return GVX_OK;
}

int gvx_robot_transform(const float* link_pts, const uint32_t* link_sizes, uint32_t n_links, const float* dh,
                        const float base[12], const float* q, int mem, float* out, int out_mem)
{
  GVX_CHECK(ensure_init());
  if (n_links == 0 || !link_sizes || !dh || !q || !out) return fail(GVX_ERR_BAD_ARG, "null argument");
  std::lock_guard<std::recursive_mutex> g(rt().mu);
  RobotStage S;
  uint64_t n = 0;
  GVX_CHECK(robot_stage(link_pts, link_sizes, n_links, dh, base, q, nullptr, 0, mem, &n, S));
  GVX_CHECK(robot_pose(S, n_links, n));
  GVX_CHECK(copy_out(out, S.pout, n * 12, out_mem));
  return sync();
}

int gvx_map_insert_robot(gvx_map map, const float* link_pts, const uint32_t* link_sizes, uint32_t n_links,
                         const float* dh, const float base[12], const float* q, const uint32_t* meanings,
                         uint32_t meaning, int mem, int* points_outside)
{
  GVX_CHECK(ensure_init());
  if (!map || !map->data) return fail(GVX_ERR_BAD_ARG, "null map");
  if (n_links == 0 || !link_sizes || !dh || !q) return fail(GVX_ERR_BAD_ARG, "null argument");
  std::lock_guard<std::recursive_mutex> g(map->mu);
  std::lock_guard<std::recursive_mutex> g2(rt().mu);
  RobotStage S;
  uint64_t n = 0;
  GVX_CHECK(robot_stage(link_pts, link_sizes, n_links, dh, base, q, meanings, meaning, mem, &n, S));
  GVX_CHECK(fill_u32(rt().dev_small, 0u, 4));
  {
    ProfScope ps("insert_robot");
    if (map->type == GVX_MT_BITVECTOR_VOXELMAP && n_links <= 512) {  // 104 B of LDS per link
      if (n) {
        hipLaunchKernelGGL(k_robot_insert_bit, dim3(grid_for(n, 256, 4096)), dim3(256), n_links * 104, rt().stream,
                           (uint32_t*)map->data, map->dims[0], map->dims[1], map->dims[2], map->side, S.pin, S.ends,
                           S.meanings, n_links, n, S.dh, S.base, S.q, (int*)rt().dev_small);
        GVX_HIP(hipGetLastError());
      }
    } else {
      GVX_CHECK(robot_pose(S, n_links, n));
      GVX_CHECK(insert_points_async(map, S.pout, n, S.ends, n_links, S.meanings, meaning, (int*)rt().dev_small));
    }
  }
  GVX_HIP(hipMemcpyAsync(rt().host_small, rt().dev_small, 4, hipMemcpyDeviceToHost, rt().stream));
  GVX_CHECK(sync());
  if (points_outside) *points_outside = *(int*)rt().host_small;
  return GVX_OK;
}

int gvx_map_insert_robot_selfcoll(gvx_map map, const float* link_pts, const uint32_t* link_sizes, uint32_t n_links,
                                  const float* dh, const float base[12], const float* q, const uint32_t* meanings,
                                  const uint8_t* masks, int mem, uint8_t colliding[32], int* self_collision,
                                  int* points_outside)
{
  GVX_CHECK(ensure_init());
  if (!map || !map->data) return fail(GVX_ERR_BAD_ARG, "null map");
  if (n_links == 0 || !link_sizes || !dh || !q) return fail(GVX_ERR_BAD_ARG, "null argument");
  std::lock_guard<std::recursive_mutex> g(map->mu);
  std::lock_guard<std::recursive_mutex> g2(rt().mu);
  RobotStage S;
  uint64_t n = 0;
  GVX_CHECK(robot_stage(link_pts, link_sizes, n_links, dh, base, q, meanings, 0, mem, &n, S));
  uint32_t* acc = (uint32_t*)rt().dev_small + 64;
  {
    ProfScope ps("insert_robot_selfcoll");
    GVX_CHECK(robot_pose(S, n_links, n));
    GVX_CHECK(selfcoll_insert_async(map, S.pout, link_sizes, n_links, meanings, masks, acc));
  }
  return selfcoll_finish(acc, colliding, self_collision, points_outside);
}

int gvx_list_insert_robot(gvx_list list, const float* link_pts, const uint32_t* link_sizes, uint32_t n_links,
                          const float* dh, const float base[12], const float* q, const uint32_t* meanings,
                          uint32_t meaning, int mem)
{
  GVX_CHECK(ensure_init());
  if (!list) return fail(GVX_ERR_BAD_ARG, "null list");
  if (n_links == 0 || !link_sizes || !dh || !q) return fail(GVX_ERR_BAD_ARG, "null argument");
  std::lock_guard<std::recursive_mutex> g2(rt().mu);
  RobotStage S;
  uint64_t n = 0;
  GVX_CHECK(robot_stage(link_pts, link_sizes, n_links, dh, base, q, meanings, meaning, mem, &n, S));
  GVX_CHECK(robot_pose(S, n_links, n));
  return list_insert_device(list, S.pout, n, S.ends, n_links, S.meanings, meaning);
}

}  // extern "C"
