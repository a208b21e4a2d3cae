// gvx_batch.hip -- batched multi-pose robot-vs-environment collision.
//
// Replaces the per-pose planner loop (clearMap, KinematicChain::setConfiguration,
// insertMetaPointCloud, collideWith; gvl_ompl_planning/gvl_ompl_planner_helper.cpp:169-200)
// by one persistent, graph-captured launch over all poses:
//   * the environment is compacted once into a bricked occupancy bitmap
//     (8x8x8-voxel bricks of 64 bytes, one cache line each) plus a per-brick
//     prefix count, so every occupied voxel has a dense rank;
//   * forward kinematics runs per pose on the GPU (deterministic sin/cos);
//   * each work item is (link chunk, group of poses): the chunk's link-frame
//     points stay in registers while the group's transforms are applied;
//   * a colliding point sets bit `pose` in row `rank` of a hit matrix with
//     atomicOr; the returned old word says whether this voxel was new for the
//     pose, which gives the exact count of distinct colliding voxels.
#include <map>

#include "gvx_internal.hpp"

namespace gvx {

int insert_points_async(MapS* m, const float* pts_dev, uint64_t n, const uint32_t* ends_dev, uint32_t ncl,
                        const uint32_t* meanings_dev, uint32_t meaning, int* outside_dev);
int map_clear_async(MapS* m);

struct Schedule {
  int block;     // points per work item (= threads per block)
  int pg;        // poses per work item (divides 32)
  int per_cu;    // resident blocks per CU of the persistent grid
};
static const Schedule kSchedules[] = {
    {256, 32, 4}, {256, 16, 4}, {256, 8, 4}, {128, 32, 8}, {128, 16, 8}, {256, 32, 8}, {512, 32, 2}, {64, 32, 16},
};
constexpr int kNumSchedules = sizeof(kSchedules) / sizeof(kSchedules[0]);

struct BatchS {
  MapS* env = nullptr;
  float thr = 1.0f;
  uint32_t dims[3];
  float side;
  uint32_t bxn, byn, bzn;
  uint64_t nbricks = 0;
  uint64_t* bricks = nullptr;   // 8 words per brick
  uint32_t* bcount = nullptr;   // popcount per brick
  uint32_t* bprefix = nullptr;  // exclusive prefix (+ total at [nbricks])
  uint32_t n_occ = 0;
  // robot
  uint32_t n_links = 0;
  uint64_t n_pts = 0;
  float* pts = nullptr;
  float* dh = nullptr;    // n_links x 5
  float* base = nullptr;  // 12
  std::vector<uint32_t> sizes;
  // chunk tables per block size: (link, start, count)
  std::map<int, std::pair<uint32_t*, uint32_t>> chunks;
  // per-call buffers
  uint32_t max_poses = 0;
  float* q = nullptr;
  float* T = nullptr;          // max_poses x n_links x 12
  uint32_t* hits = nullptr;    // n_occ x words
  uint64_t hits_words = 0;
  uint32_t* counts = nullptr;  // max_poses
  uint32_t* masks = nullptr;   // max_poses x 8
  int schedule = 0;
  std::map<std::pair<uint32_t, int>, hipGraphExec_t> graphs;
  std::recursive_mutex mu;
};

}  // namespace gvx

struct gvx_batch_s : gvx::BatchS {};

namespace gvx {

// ---------------------------------------------------------------- environment bitmap
// one wave per brick: lane (z_l * 8 + y_l) assembles the 8-voxel x row into byte lane
template <int T>
__global__ void __launch_bounds__(256) k_env_bricks(const uint8_t* __restrict__ env, uint32_t dx, uint32_t dy,
                                                    uint32_t dz, uint32_t bxn, uint32_t byn, uint64_t nbricks, int thr,
                                                    uint8_t* __restrict__ bricks, uint32_t* __restrict__ bcount)
{
  const uint64_t brick = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (brick >= nbricks) return;
  const int lane = threadIdx.x & 63;
  const uint32_t bx = (uint32_t)(brick % bxn), by = (uint32_t)((brick / bxn) % byn), bz = (uint32_t)(brick / ((uint64_t)bxn * byn));
  const uint32_t y = by * 8 + (lane & 7), z = bz * 8 + (lane >> 3);
  uint32_t row = 0;
  if (y < dy && z < dz) {
#pragma unroll
    for (int xl = 0; xl < 8; ++xl) {
      const uint32_t x = bx * 8 + xl;
      if (x < dx) {
        const uint64_t i = ((uint64_t)z * dy + y) * dx + x;
        bool occ;
        if constexpr (T == GVX_MT_BITVECTOR_VOXELMAP) {
          const uint4* p = reinterpret_cast<const uint4*>(env + i * 32);
          occ = bits_occupied(p[0], p[1]);
        } else {
          occ = (int)(int8_t)env[i] >= thr;
        }
        row |= (occ ? 1u : 0u) << xl;
      }
    }
  }
  // byte index z_l * 8 + y_l: word z_l (64 bits) holds bit (y_l * 8 + x_l)
  bricks[brick * 64 + (uint64_t)((lane >> 3) * 8 + (lane & 7))] = (uint8_t)row;
  uint32_t c = (uint32_t)__popc(row);
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if (lane == 0) bcount[brick] = c;
}

// ---------------------------------------------------------------- forward kinematics
// KinematicChain::setConfiguration (robot/dh_robot/KinematicChain.cu:93-126):
// link l is posed by base * A_0(q_0) ... A_{l-1}(q_{l-1})
__global__ void k_batch_fk(const float* __restrict__ dh, const float* __restrict__ base, const float* __restrict__ q,
                           uint32_t n_poses, uint32_t nl, float* __restrict__ T)
{
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_poses) return;
  float t[12], a[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) t[k] = base[k];
  for (uint32_t l = 0; l < nl; ++l) {
    float* o = T + ((uint64_t)p * nl + l) * 12;
#pragma unroll
    for (int k = 0; k < 12; ++k) o[k] = t[k];
    dh_matrix(dh + 5 * l, q[(uint64_t)p * nl + l], a);
    affine_mul(t, a, t);
  }
}

// ---------------------------------------------------------------- main kernel
template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_batch_collide(const float* __restrict__ pts,
                                                        const uint32_t* __restrict__ chunks, uint32_t n_chunks,
                                                        const float* __restrict__ T, uint32_t nl, uint32_t n_poses,
                                                        uint32_t pg, const uint64_t* __restrict__ bricks,
                                                        const uint32_t* __restrict__ bprefix, uint32_t dx, uint32_t dy,
                                                        uint32_t dz, uint32_t bxn, uint32_t byn, float side,
                                                        uint32_t* __restrict__ hits, uint32_t words,
                                                        uint32_t* __restrict__ counts, uint32_t* __restrict__ masks)
{
  __shared__ float sT[32 * 12];
  const uint32_t groups = (n_poses + pg - 1) / pg;
  const uint64_t items = (uint64_t)n_chunks * groups;
  const int lane = threadIdx.x & 63;
  for (uint64_t it = blockIdx.x; it < items; it += gridDim.x) {
    const uint32_t c = (uint32_t)(it % n_chunks), g = (uint32_t)(it / n_chunks);
    const uint32_t link = chunks[3 * c], start = chunks[3 * c + 1], cnt = chunks[3 * c + 2];
    const uint32_t p0 = g * pg;
    const uint32_t np = (n_poses - p0) < pg ? (n_poses - p0) : pg;
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < np * 12; k += BLOCK) sT[k] = T[((uint64_t)(p0 + k / 12) * nl + link) * 12 + k % 12];
    __syncthreads();
    const bool valid = threadIdx.x < cnt;
    float px = 0.f, py = 0.f, pz = 0.f;
    if (valid) {
      const uint64_t i = (uint64_t)start + threadIdx.x;
      px = pts[3 * i]; py = pts[3 * i + 1]; pz = pts[3 * i + 2];
    }
    const uint32_t mbit = 1u << ((M_SV_START + link) & 31);
    const uint32_t mword = (M_SV_START + link) >> 5;
    for (uint32_t j = 0; j < np; ++j) {
      const uint32_t pose = p0 + j;
      bool hit = false, isnew = false;
      if (valid) {
        float wx, wy, wz;
        transform(sT + 12 * j, px, py, pz, wx, wy, wz);
        const uint32_t x = quantize(wx, side), y = quantize(wy, side), z = quantize(wz, side);
        if (x < dx && y < dy && z < dz) {
          const uint32_t brick = ((z >> 3) * byn + (y >> 3)) * bxn + (x >> 3);  // bricks of a map in memory: < 2^32
          const uint32_t zl = z & 7, bit = ((y & 7) << 3) | (x & 7);
          const uint64_t* bw = bricks + brick * 8;
          const uint64_t w = bw[zl];
          if ((w >> bit) & 1ull) {
            hit = true;
            uint32_t rank = bprefix[brick] + (uint32_t)__popcll(w & ((1ull << bit) - 1ull));
            for (uint32_t k = 0; k < zl; ++k) rank += (uint32_t)__popcll(bw[k]);
            const uint32_t b = 1u << (pose & 31);
            const uint32_t old = atomicOr(hits + (uint64_t)rank * words + (pose >> 5), b);
            isnew = (old & b) == 0u;
          }
        }
      }
      const uint64_t nb = wave_ballot(isnew);
      const uint64_t hb = wave_ballot(hit);
      if (lane == 0) {
        if (nb) atomicAdd(counts + pose, (uint32_t)__popcll(nb));
        if (hb) atomicOr(masks + (uint64_t)pose * 8 + mword, mbit);
      }
    }
  }
}

static int batch_rebuild_env(BatchS* b)
{
  MapS* e = b->env;
  hipStream_t s = rt().stream;
  const int thr = float_to_probability(b->thr);
  unsigned g = (unsigned)((b->nbricks + 3) / 4);
  if (e->type == GVX_MT_BITVECTOR_VOXELMAP)
    hipLaunchKernelGGL(k_env_bricks<GVX_MT_BITVECTOR_VOXELMAP>, dim3(g), dim3(256), 0, s, (const uint8_t*)e->data,
                       b->dims[0], b->dims[1], b->dims[2], b->bxn, b->byn, b->nbricks, thr, (uint8_t*)b->bricks,
                       b->bcount);
  else
    hipLaunchKernelGGL(k_env_bricks<GVX_MT_PROBAB_VOXELMAP>, dim3(g), dim3(256), 0, s, (const uint8_t*)e->data,
                       b->dims[0], b->dims[1], b->dims[2], b->bxn, b->byn, b->nbricks, thr, (uint8_t*)b->bricks,
                       b->bcount);
  GVX_HIP(hipGetLastError());
  GVX_CHECK(scan_u32(b->bcount, b->bprefix, b->nbricks, b->bprefix + b->nbricks));
  GVX_HIP(hipMemcpyAsync(rt().host_small, b->bprefix + b->nbricks, 4, hipMemcpyDeviceToHost, s));
  GVX_CHECK(sync());
  b->n_occ = *(uint32_t*)rt().host_small;
  const uint64_t words = (b->max_poses + 31) / 32;
  const uint64_t need = (uint64_t)(b->n_occ ? b->n_occ : 1) * words;
  // captured graphs hold the hit-matrix size of the previous environment
  for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
  b->graphs.clear();
  if (need > b->hits_words) {
    dfree(b->hits);
    b->hits = nullptr;
    b->hits_words = 0;
    GVX_CHECK(dalloc((void**)&b->hits, need * 4));
    b->hits_words = need;
  }
  return GVX_OK;
}

static int chunk_table(BatchS* b, int block, uint32_t** tab, uint32_t* n)
{
  auto it = b->chunks.find(block);
  if (it != b->chunks.end()) {
    *tab = it->second.first;
    *n = it->second.second;
    return GVX_OK;
  }
  std::vector<uint32_t> t;
  uint64_t off = 0;
  for (uint32_t l = 0; l < b->n_links; ++l) {
    for (uint32_t s = 0; s < b->sizes[l]; s += block) {
      t.push_back(l);
      t.push_back((uint32_t)(off + s));
      t.push_back(b->sizes[l] - s < (uint32_t)block ? b->sizes[l] - s : (uint32_t)block);
    }
    off += b->sizes[l];
  }
  uint32_t* d = nullptr;
  GVX_CHECK(dalloc((void**)&d, t.size() * 4 + 4));
  if (!t.empty()) GVX_HIP(hipMemcpyAsync(d, t.data(), t.size() * 4, hipMemcpyHostToDevice, rt().stream));
  GVX_CHECK(sync());
  b->chunks[block] = {d, (uint32_t)(t.size() / 3)};
  *tab = d;
  *n = (uint32_t)(t.size() / 3);
  return GVX_OK;
}

// enqueue one batched evaluation of n poses (q already in b->q)
static int batch_enqueue(BatchS* b, uint32_t n, int sched)
{
  hipStream_t s = rt().stream;
  const Schedule& S = kSchedules[sched];
  uint32_t* tab = nullptr;
  uint32_t nch = 0;
  GVX_CHECK(chunk_table(b, S.block, &tab, &nch));
  const uint32_t words = (b->max_poses + 31) / 32;
  GVX_CHECK(fill_u32(b->hits, 0u, (uint64_t)(b->n_occ ? b->n_occ : 1) * words));
  GVX_CHECK(fill_u32(b->counts, 0u, n));
  GVX_CHECK(fill_u32(b->masks, 0u, (uint64_t)n * 8));
  hipLaunchKernelGGL(k_batch_fk, dim3((n + 63) / 64), dim3(64), 0, s, b->dh, b->base, b->q, n, b->n_links, b->T);
  GVX_HIP(hipGetLastError());
  if (nch == 0 || b->n_occ == 0) return GVX_OK;
  const uint64_t items = (uint64_t)nch * ((n + S.pg - 1) / S.pg);
  uint64_t grid = (uint64_t)(rt().num_cus > 0 ? rt().num_cus : 256) * S.per_cu;
  if (grid > items) grid = items;
#define LAUNCH(BS)                                                                                                     \
  hipLaunchKernelGGL(k_batch_collide<BS>, dim3((unsigned)grid), dim3(BS), 0, s, b->pts, tab, nch, b->T, b->n_links, n, \
                     (uint32_t)S.pg, b->bricks, b->bprefix, b->dims[0], b->dims[1], b->dims[2], b->bxn, b->byn,      \
                     b->side, b->hits, words, b->counts, b->masks)
  switch (S.block) {
    case 64: LAUNCH(64); break;
    case 128: LAUNCH(128); break;
    case 256: LAUNCH(256); break;
    default: LAUNCH(512); break;
  }
#undef LAUNCH
  GVX_HIP(hipGetLastError());
  return GVX_OK;
}

static int batch_run(BatchS* b, uint32_t n, int sched, bool use_graph)
{
  hipStream_t s = rt().stream;
  if (!use_graph) return batch_enqueue(b, n, sched);
  auto key = std::make_pair(n, sched);
  auto it = b->graphs.find(key);
  if (it == b->graphs.end()) {
    // make sure lazily built tables exist before capture
    uint32_t* tab;
    uint32_t nch;
    GVX_CHECK(chunk_table(b, kSchedules[sched].block, &tab, &nch));
    hipGraph_t graph;
    GVX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int st = batch_enqueue(b, n, sched);
    hipError_t e = hipStreamEndCapture(s, &graph);
    if (st != GVX_OK) return st;
    if (e != hipSuccess) return fail(GVX_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    hipGraphExec_t exec;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(GVX_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    it = b->graphs.emplace(key, exec).first;
  }
  ProfScope ps("batch_collide_graph");
  GVX_HIP(hipGraphLaunch(it->second, s));
  return GVX_OK;
}

// ---------------------------------------------------------------- per-pose reference path
// kernelTransformCloud fused with kernelInsertMetaPointCloud: link l posed by T[l], meaning 4 + l
__global__ void __launch_bounds__(256) k_insert_posed(uint32_t* __restrict__ map, uint32_t dx, uint32_t dy, uint32_t dz,
                                                      float side, const float* __restrict__ pts,
                                                      const uint32_t* __restrict__ ends, uint32_t nl, uint64_t n,
                                                      const float* __restrict__ T)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    uint32_t lo = 0, hi = nl - 1;
    while (lo < hi) {
      uint32_t mid = (lo + hi) >> 1;
      if ((uint64_t)ends[mid] > i) hi = mid; else lo = mid + 1;
    }
    float wx, wy, wz;
    transform(T + 12 * lo, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], wx, wy, wz);
    const uint32_t x = quantize(wx, side), y = quantize(wy, side), z = quantize(wz, side);
    if (x < dx && y < dy && z < dz) {
      const uint32_t m = M_SV_START + lo;
      atomicOr(map + (((uint64_t)z * dy + y) * dx + x) * 8 + (m >> 5), 1u << (m & 31));
    }
  }
}

// robot (BitVector) x env collideWith count, DefaultCollider, result added to out[slot]
template <int TB>
__global__ void __launch_bounds__(256) k_pose_collide(uint8_t* __restrict__ a, const uint8_t* __restrict__ e, uint64_t n,
                                                      int thr, uint32_t* __restrict__ out)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint32_t c = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    bool occ_e;
    if constexpr (TB == GVX_MT_BITVECTOR_VOXELMAP) {
      const uint4* p = reinterpret_cast<const uint4*>(e + i * 32);
      occ_e = bits_occupied(p[0], p[1]);
    } else {
      occ_e = (int)(int8_t)e[i] >= thr;
    }
    bool hit = false;
    if (occ_e) {
      uint4* pa = reinterpret_cast<uint4*>(a + i * 32);
      uint4 a0 = pa[0], a1 = pa[1];
      hit = bits_occupied(a0, a1);
      if (hit) reinterpret_cast<uint32_t*>(pa)[0] = a0.x | (1u << M_COLLISION);
    }
    c += (uint32_t)__popcll(wave_ballot(hit));
  }
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(out, c);
}

}  // namespace gvx

using namespace gvx;

extern "C" {

int gvx_batch_create(gvx_map env, float threshold, const float* link_pts, const uint32_t* link_sizes, uint32_t n_links,
                     const float* dh, const float base[12], uint32_t max_poses, gvx_batch* out)
{
  if (!out) return fail(GVX_ERR_BAD_ARG, "null out");
  *out = nullptr;
  GVX_CHECK(ensure_init());
  if (!env || !env->data) return fail(GVX_ERR_BAD_ARG, "null env map");
  if (env->type != GVX_MT_BITVECTOR_VOXELMAP && env->type != GVX_MT_PROBAB_VOXELMAP)
    return fail(GVX_ERR_UNSUPPORTED, "environment must be a BitVector or Prob map");
  if (n_links == 0 || n_links > 250 || !link_sizes || !dh) return fail(GVX_ERR_BAD_ARG, "bad robot description");
  if (max_poses == 0) return fail(GVX_ERR_BAD_ARG, "max_poses == 0");
  gvx_batch b = new gvx_batch_s();
  auto bail = [&](int s) {
    gvx_batch_destroy(b);
    return s;
  };
  b->env = env;
  b->thr = threshold;
  memcpy(b->dims, env->dims, sizeof(b->dims));
  b->side = env->side;
  b->bxn = (b->dims[0] + 7) / 8;
  b->byn = (b->dims[1] + 7) / 8;
  b->bzn = (b->dims[2] + 7) / 8;
  b->nbricks = (uint64_t)b->bxn * b->byn * b->bzn;
  b->n_links = n_links;
  b->sizes.assign(link_sizes, link_sizes + n_links);
  uint64_t tot = 0;
  for (uint32_t l = 0; l < n_links; ++l) tot += link_sizes[l];
  if (tot > 0xFFFFFFFFull) return bail(fail(GVX_ERR_UNSUPPORTED, "too many link points"));
  if (tot && !link_pts) return bail(fail(GVX_ERR_BAD_ARG, "null link points"));
  b->n_pts = tot;
  b->max_poses = max_poses;
  int s;
  if ((s = dalloc((void**)&b->bricks, b->nbricks * 64)) != GVX_OK) return bail(s);
  if ((s = dalloc((void**)&b->bcount, b->nbricks * 4)) != GVX_OK) return bail(s);
  if ((s = dalloc((void**)&b->bprefix, (b->nbricks + 1) * 4)) != GVX_OK) return bail(s);
  if ((s = dalloc((void**)&b->pts, tot * 12 + 16)) != GVX_OK) return bail(s);
  if ((s = dalloc((void**)&b->dh, n_links * 20)) != GVX_OK) return bail(s);
  if ((s = dalloc((void**)&b->base, 48)) != GVX_OK) return bail(s);
  if ((s = dalloc((void**)&b->q, (uint64_t)max_poses * n_links * 4)) != GVX_OK) return bail(s);
  if ((s = dalloc((void**)&b->T, (uint64_t)max_poses * n_links * 48)) != GVX_OK) return bail(s);
  if ((s = dalloc((void**)&b->counts, (uint64_t)max_poses * 4)) != GVX_OK) return bail(s);
  if ((s = dalloc((void**)&b->masks, (uint64_t)max_poses * 32)) != GVX_OK) return bail(s);
  float ident[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
  hipStream_t st = rt().stream;
  if (tot && hipMemcpyAsync(b->pts, link_pts, tot * 12, hipMemcpyHostToDevice, st) != hipSuccess)
    return bail(fail(GVX_ERR_HIP, "copy link points"));
  if (hipMemcpyAsync(b->dh, dh, n_links * 20, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(b->base, base ? base : ident, 48, hipMemcpyHostToDevice, st) != hipSuccess)
    return bail(fail(GVX_ERR_HIP, "copy robot description"));
  {
    std::lock_guard<std::recursive_mutex> g(env->mu);
    std::lock_guard<std::recursive_mutex> g2(rt().mu);
    if ((s = batch_rebuild_env(b)) != GVX_OK) return bail(s);
  }
  *out = b;
  return GVX_OK;
}

int gvx_batch_destroy(gvx_batch b)
{
  if (!b) return GVX_OK;
  (void)sync();
  for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto& kv : b->chunks) dfree(kv.second.first);
  void* bufs[] = {b->bricks, b->bcount, b->bprefix, b->pts, b->dh, b->base, b->q, b->T, b->hits, b->counts, b->masks};
  for (auto p : bufs) dfree(p);
  delete b;
  return GVX_OK;
}

int gvx_batch_refresh_env(gvx_batch b)
{
  if (!b) return fail(GVX_ERR_BAD_ARG, "null batch");
  std::lock_guard<std::recursive_mutex> g0(b->mu);
  std::lock_guard<std::recursive_mutex> g(b->env->mu);
  std::lock_guard<std::recursive_mutex> g2(rt().mu);
  return batch_rebuild_env(b);
}

static int batch_collide_impl(gvx_batch b, const float* q, uint32_t n, int mem, uint32_t* counts, uint8_t* masks,
                              int out_mem, int sched, bool use_graph)
{
  if (!b) return fail(GVX_ERR_BAD_ARG, "null batch");
  if (n > b->max_poses) return fail(GVX_ERR_BAD_ARG, "n_poses > max_poses");
  if (n && !q) return fail(GVX_ERR_BAD_ARG, "null q");
  if (n == 0) return GVX_OK;
  std::lock_guard<std::recursive_mutex> g0(b->mu);
  std::lock_guard<std::recursive_mutex> g(b->env->mu);
  std::lock_guard<std::recursive_mutex> g2(rt().mu);
  GVX_CHECK(copy_in(b->q, q, (uint64_t)n * b->n_links * 4, mem));
  GVX_CHECK(batch_run(b, n, sched, use_graph));
  if (counts) GVX_CHECK(copy_out(counts, b->counts, (uint64_t)n * 4, out_mem));
  if (masks) GVX_CHECK(copy_out(masks, b->masks, (uint64_t)n * 32, out_mem));
  return sync();
}

int gvx_batch_collide(gvx_batch b, const float* q, uint32_t n, int mem, uint32_t* counts, uint8_t* masks, int out_mem)
{
  if (!b) return fail(GVX_ERR_BAD_ARG, "null batch");
  return batch_collide_impl(b, q, n, mem, counts, masks, out_mem, b->schedule, true);
}

int gvx_batch_set_schedule(gvx_batch b, int schedule)
{
  if (!b) return fail(GVX_ERR_BAD_ARG, "null batch");
  if (schedule < 0 || schedule >= kNumSchedules) return fail(GVX_ERR_BAD_ARG, "schedule out of range");
  b->schedule = schedule;
  return GVX_OK;
}

int gvx_batch_autotune(gvx_batch b, uint32_t n, int* schedule_out)
{
  if (!b) return fail(GVX_ERR_BAD_ARG, "null batch");
  if (n == 0 || n > b->max_poses) return fail(GVX_ERR_BAD_ARG, "bad n_poses");
  std::lock_guard<std::recursive_mutex> g0(b->mu);
  std::lock_guard<std::recursive_mutex> g(b->env->mu);
  std::lock_guard<std::recursive_mutex> g2(rt().mu);
  // tune on the poses currently in b->q (the last call's), or zeros
  hipEvent_t e0, e1;
  GVX_HIP(hipEventCreate(&e0));
  GVX_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  int best_s = b->schedule;
  for (int s = 0; s < kNumSchedules; ++s) {
    GVX_CHECK(batch_run(b, n, s, true));  // warm-up + capture
    GVX_HIP(hipEventRecord(e0, rt().stream));
    for (int r = 0; r < 3; ++r) GVX_CHECK(batch_run(b, n, s, true));
    GVX_HIP(hipEventRecord(e1, rt().stream));
    GVX_CHECK(sync());
    float ms = 0.f;
    GVX_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best) { best = ms; best_s = s; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->schedule = best_s;
  if (schedule_out) *schedule_out = best_s;
  return GVX_OK;
}

int gvx_batch_collide_per_pose(gvx_batch b, gvx_map robot, const float* q, uint32_t n, int mem, uint32_t* counts)
{
  if (!b || !robot || !robot->data) return fail(GVX_ERR_BAD_ARG, "null batch/robot");
  if (robot->type != GVX_MT_BITVECTOR_VOXELMAP) return fail(GVX_ERR_UNSUPPORTED, "robot map must be a BitVector map");
  if (memcmp(robot->dims, b->dims, sizeof(b->dims)) != 0) return fail(GVX_ERR_DIM_MISMATCH, "robot/env dims differ");
  if (n > b->max_poses) return fail(GVX_ERR_BAD_ARG, "n_poses > max_poses");
  if (n == 0) return GVX_OK;
  std::lock_guard<std::recursive_mutex> g0(b->mu);
  std::unique_lock<std::recursive_mutex> g1(robot->mu, std::defer_lock), g2(b->env->mu, std::defer_lock);
  std::lock(g1, g2);
  std::lock_guard<std::recursive_mutex> g3(rt().mu);
  hipStream_t s = rt().stream;
  GVX_CHECK(copy_in(b->q, q, (uint64_t)n * b->n_links * 4, mem));
  // link end offsets for the fused transform + insert
  std::vector<uint32_t> ends(b->n_links);
  uint64_t acc = 0;
  for (uint32_t l = 0; l < b->n_links; ++l) { acc += b->sizes[l]; ends[l] = (uint32_t)acc; }
  void* st = nullptr;
  GVX_CHECK(stage_reserve(ends.size() * 4, &st));
  GVX_HIP(hipMemcpyAsync(st, ends.data(), ends.size() * 4, hipMemcpyHostToDevice, s));
  GVX_CHECK(fill_u32(b->counts, 0u, n));
  hipLaunchKernelGGL(k_batch_fk, dim3((n + 63) / 64), dim3(64), 0, s, b->dh, b->base, b->q, n, b->n_links, b->T);
  GVX_HIP(hipGetLastError());
  const int thr = float_to_probability(b->thr);
  const unsigned gcol = grid_for(robot->n, 256, (unsigned)(rt().num_cus > 0 ? rt().num_cus * 8 : 2048));
  for (uint32_t p = 0; p < n; ++p) {
    ProfScope ps("per_pose_collide");
    GVX_CHECK(map_clear_async(robot));
    if (b->n_pts)
      hipLaunchKernelGGL(k_insert_posed, dim3(grid_for(b->n_pts, 256, 1u << 14)), dim3(256), 0, s,
                         (uint32_t*)robot->data, b->dims[0], b->dims[1], b->dims[2], b->side, b->pts,
                         (const uint32_t*)st, b->n_links, b->n_pts, b->T + (uint64_t)p * b->n_links * 12);
    if (b->env->type == GVX_MT_BITVECTOR_VOXELMAP)
      hipLaunchKernelGGL(k_pose_collide<GVX_MT_BITVECTOR_VOXELMAP>, dim3(gcol), dim3(256), 0, s,
                         (uint8_t*)robot->data, (const uint8_t*)b->env->data, robot->n, thr, b->counts + p);
    else
      hipLaunchKernelGGL(k_pose_collide<GVX_MT_PROBAB_VOXELMAP>, dim3(gcol), dim3(256), 0, s, (uint8_t*)robot->data,
                         (const uint8_t*)b->env->data, robot->n, thr, b->counts + p);
    GVX_HIP(hipGetLastError());
  }
  if (counts) GVX_CHECK(copy_out(counts, b->counts, (uint64_t)n * 4, GVX_MEM_HOST));
  return sync();
}

}  // extern "C"
