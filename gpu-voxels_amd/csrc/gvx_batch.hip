// gvx_batch.hip -- batched multi-pose robot-vs-environment collision.
//
// Replaces the per-pose planner loop (clearMap, KinematicChain::setConfiguration,
// insertMetaPointCloud, collideWith; gvl_ompl_planning/gvl_ompl_planner_helper.cpp:169-200)
// by one graph-captured launch over all poses:
//   * the environment is compacted once into a bricked occupancy bitmap
//     (8x8x8-voxel bricks of 64 bytes, one cache line each) plus a per-brick
//     prefix count, so every occupied voxel has a dense rank, and a one-bit-per-brick
//     occupancy summary;
//   * each link's points are stored along a Morton curve of the link frame, so a
//     chunk of consecutive points covers a small patch with a tight bounding box;
//   * forward kinematics runs per pose on the GPU (deterministic sin/cos);
//   * each work item is (chunk, group of poses): one thread per pose first poses the
//     chunk's box and drops the poses whose box reaches no occupied brick (a
//     conservative test, so no colliding point is ever dropped), then the chunk's
//     points, one per thread, are tested against the surviving poses;
//   * a colliding point sets bit `pose` in row `rank` of a hit matrix with
//     atomicOr (in dense environments once per run of consecutive lanes on one
//     voxel); the returned old word says whether this voxel was new for the pose,
//     which gives the exact count of distinct colliding voxels;
//   * the hit matrix is never cleared as a whole: a word that goes from zero to
//     non-zero is appended to the touched-word list of its slot (word index mod
//     256; a slot can hold every word of its residue, so it never overflows), and
//     the next call's prologue zeroes exactly those words, one block per slot, so a
//     call costs what its collisions cost, whatever the number of occupied
//     environment voxels.
#include <algorithm>
#include <map>
#include <tuple>

#include "gvx_internal.hpp"

namespace gvx {

int insert_points_async(MapS* m, const float* pts_dev, uint64_t n, const uint32_t* ends_dev, uint32_t ncl,
                        const uint32_t* meanings_dev, uint32_t meaning, int* outside_dev);
int map_clear_async(MapS* m);

struct Schedule {
  int chunk;    // points per chunk
  int block;    // threads per block
  int ppt;      // poses per thread in the broad phase: block * ppt poses per work item
  int persist;  // 0: one block per work item; k: k blocks per CU striding over the items
};
static const Schedule kSchedules[] = {
    {128, 512, 1, 0}, {128, 256, 1, 0}, {64, 256, 1, 0}, {256, 256, 1, 0},
    {128, 512, 1, 4}, {128, 256, 1, 8}, {64, 256, 1, 8}, {128, 256, 2, 0},
};
constexpr int kKernelSchedules = sizeof(kSchedules) / sizeof(kSchedules[0]);
// schedule index: kernel schedule (index % kKernelSchedules), launched directly (index
// < kKernelSchedules) or as a replay of a captured HIP graph (the others)
constexpr int kNumSchedules = 2 * kKernelSchedules;
// broad-phase statistics are summed in this many slots, one 128-byte line each, so
// that blocks do not serialise on one address
constexpr uint32_t kStatSlots = 256, kStatStride = 16;
// touched hit-matrix words are listed in this many slot lists (slot = word index
// mod kTouchSlots, entry = word index / kTouchSlots), each slot's count on its own
// 64-byte line
constexpr uint32_t kTouchSlots = 256, kTouchStride = 16;

struct ChunkTab {
  uint32_t* tab = nullptr;  // (link, start, count) per chunk
  float* box = nullptr;     // link-frame bounding box (lo xyz, hi xyz) per chunk
  uint32_t n = 0;
};

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
  uint32_t* wrank = nullptr;    // per brick word: rank of its first occupied voxel (bprefix + earlier words)
  uint32_t* bsat = nullptr;     // summed-area table of occupied bricks, (bxn+1) x (byn+1) x (bzn+1)
  uint32_t n_occ = 0;
  // robot
  uint32_t n_links = 0;
  uint64_t n_pts = 0;
  float* pts = nullptr;
  float* dh = nullptr;    // n_links x 5
  float* base = nullptr;  // 12
  std::vector<uint32_t> sizes;
  std::vector<float> hpts;  // host copy of pts (Morton order within each link)
  float* lbox = nullptr;    // link-frame bounding box per link (lo xyz, hi xyz; NaN: a point is not finite)
  uint32_t* llive = nullptr;  // n_links x lwords: pose bit set when the link's box may reach an occupied brick
  uint32_t lwords = 0;
  // chunk tables per chunk size
  std::map<int, ChunkTab> chunks;
  // per-call buffers
  uint32_t max_poses = 0;
  float* q = nullptr;
  float* T = nullptr;          // max_poses x n_links x 12
  uint32_t* hits = nullptr;    // n_occ x words, all zero between calls but for the listed words
  uint64_t hits_words = 0;
  uint32_t* tlist = nullptr;  // kTouchSlots lists of tcap entries
  uint32_t* tcnt = nullptr;   // per slot: entries appended by the last call
  uint32_t tcap = 0;
  uint32_t* counts = nullptr;  // max_poses
  uint32_t* masks = nullptr;   // max_poses x 8
  unsigned long long* stats = nullptr;  // last call, per slot: surviving (chunk, pose) pairs, their point-poses
  uint64_t last_tests = 0;              // last call: (chunk, pose) pairs tested
  int schedule = 0;
  // captured calls: (poses, schedule, joint values, counts, masks) -> graph
  std::map<std::tuple<uint32_t, int, const void*, void*, void*>, hipGraphExec_t> graphs;
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

// ---------------------------------------------------------------- brick summed-area table
// sat[(z * (byn + 1) + y) * (bxn + 1) + x] = number of bricks holding an occupied voxel
// with bz < z, by < y, bx < x: the occupied-brick count of any box is 8 lookups.
__global__ void k_sat_init(const uint32_t* __restrict__ bcount, uint32_t bxn, uint32_t byn, uint32_t bzn,
                           uint32_t* __restrict__ sat)
{
  const uint64_t n = (uint64_t)(bxn + 1) * (byn + 1) * (bzn + 1);
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t x = (uint32_t)(i % (bxn + 1)), y = (uint32_t)((i / (bxn + 1)) % (byn + 1)),
                 z = (uint32_t)(i / ((uint64_t)(bxn + 1) * (byn + 1)));
  uint32_t v = 0;
  if (x && y && z) v = bcount[((uint64_t)(z - 1) * byn + (y - 1)) * bxn + (x - 1)] ? 1u : 0u;
  sat[i] = v;
}

// inclusive running sum along one axis (stride st, length len) for every line
__global__ void k_sat_scan(uint32_t* __restrict__ sat, uint32_t ex, uint32_t ey, uint32_t ez, int axis)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t base, st;
  uint32_t len;
  if (axis == 0) {  // lines along x: one per (y, z)
    if (i >= (uint64_t)ey * ez) return;
    base = i * ex; st = 1; len = ex;
  } else if (axis == 1) {  // along y: one per (x, z)
    if (i >= (uint64_t)ex * ez) return;
    base = (i / ex) * ex * ey + i % ex; st = ex; len = ey;
  } else {  // along z: one per (x, y)
    if (i >= (uint64_t)ex * ey) return;
    base = i; st = (uint64_t)ex * ey; len = ez;
  }
  uint32_t acc = 0;
  for (uint32_t k = 0; k < len; ++k) {
    acc += sat[base + k * st];
    sat[base + k * st] = acc;
  }
}

// wrank[brick * 8 + k] = bprefix[brick] + occupied voxels in words 0 .. k-1 of the brick
__global__ void k_word_rank(const uint64_t* __restrict__ bricks, const uint32_t* __restrict__ bprefix,
                            uint64_t nbricks, uint32_t* __restrict__ wrank)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nbricks) return;
  uint32_t r = bprefix[i];
  for (int k = 0; k < 8; ++k) {
    wrank[i * 8 + k] = r;
    r += (uint32_t)__popcll(bricks[i * 8 + k]);
  }
}

// Can a point of the link-frame box bx = (lo xyz, hi xyz) posed by the 3x4 affine m
// land in a brick holding an occupied voxel?  false only when no point of the box
// can.  Per axis the box's image lies in [mn, mx] (interval arithmetic on m's row);
// the range is widened by two voxel sides plus 1e-5 of the row's magnitude before it
// is scaled by 1/side, which covers the roundings of transform() (three, each below
// 2^-24 of that magnitude), of this bound and of the reciprocal scaling, so the
// voxel of every point, floor(p / side) saturated to [0, 2^32), lies between the two
// saturated ends.  Non-finite or huge (> 1e30) magnitudes answer true.
__device__ __forceinline__ bool box_may_hit(const float* m, const float* bx, float side, float inv_side,
                                            uint32_t dx, uint32_t dy, uint32_t dz, const uint32_t* __restrict__ sat,
                                            uint32_t bxn, uint32_t byn)
{
  uint32_t v0[3], v1[3];
  const uint32_t dim[3] = {dx, dy, dz};
  bool big = false, out = false;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    float mn = m[4 * r + 3], mx = mn, mag = fabsf(mn);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float a = m[4 * r + k] * bx[k], b = m[4 * r + k] * bx[3 + k];
      mn += fminf(a, b);
      mx += fmaxf(a, b);
      mag += fabsf(m[4 * r + k]) * (fabsf(bx[k]) + fabsf(bx[3 + k]));  // NaN and inf propagate
    }
    big = big || !(mag <= 1e30f);
    const float e = 2.0f * side + 1e-5f * mag;
    const float f0 = fminf(fmaxf(floorf((mn - e) * inv_side), 0.0f), 4294967040.0f);
    const float f1 = fminf(fmaxf(floorf((mx + e) * inv_side), 0.0f), 4294967040.0f);
    v0[r] = (uint32_t)f0;
    v1[r] = (uint32_t)f1;
    out = out || v0[r] >= dim[r];  // every point of the box is outside the map on this axis
    v0[r] = v0[r] < dim[r] - 1 ? v0[r] : dim[r] - 1;  // in range for the lookups, which are made
    v1[r] = v1[r] < dim[r] - 1 ? v1[r] : dim[r] - 1;  // unconditionally so that they can overlap
  }
  const uint64_t sy = bxn + 1, sz = (uint64_t)(bxn + 1) * (byn + 1);
  const uint64_t x0 = v0[0] >> 3, x1 = (v1[0] >> 3) + 1, y0 = (v0[1] >> 3) * sy, y1 = ((v1[1] >> 3) + 1) * sy,
                 z0 = (v0[2] >> 3) * sz, z1 = ((v1[2] >> 3) + 1) * sz;
  const uint32_t n = sat[z1 + y1 + x1] - sat[z0 + y1 + x1] - sat[z1 + y0 + x1] - sat[z1 + y1 + x0] +
                     sat[z0 + y0 + x1] + sat[z0 + y1 + x0] + sat[z1 + y0 + x0] - sat[z0 + y0 + x0];
  return big || (!out && n != 0u);
}

// ---------------------------------------------------------------- prologue
// Blocks [0, fk_blocks) own 32 poses each: the DH matrices of every (pose, link) pair
// are built in parallel, one thread per pose chains them (KinematicChain::setConfiguration,
// robot/dh_robot/KinematicChain.cu:93-126: link l is posed by base * A_0 ... A_{l-1}),
// every (pose, link) pair tests the whole link's box against the summed-area table,
// and the block writes word p/32 of each link's live-pose mask, zeroes its poses'
// counts and masks.  The other blocks zero the hit-matrix words the previous call
// listed, each block the lists of its own slots, and reset those lists' counts.
constexpr int FK_SEG = 16;  // links per LDS segment
__global__ void __launch_bounds__(256) k_batch_prologue(
    uint32_t* __restrict__ hits, const uint32_t* __restrict__ tlist, uint32_t* __restrict__ tcnt, uint32_t tcap,
    uint32_t* __restrict__ counts, uint32_t* __restrict__ masks,
    const float* __restrict__ dh, const float* __restrict__ base, const float* __restrict__ q, uint32_t n_poses,
    uint32_t nl, float* __restrict__ T, unsigned long long* __restrict__ stats, const float* __restrict__ lbox,
    uint32_t* __restrict__ llive, uint32_t lwords, uint32_t fk_blocks, const uint32_t* __restrict__ sat, uint32_t dx,
    uint32_t dy, uint32_t dz, uint32_t bxn, uint32_t byn, float side, float inv_side)
{
  const uint32_t tid = threadIdx.x;
  if (blockIdx.x < fk_blocks) {
    __shared__ float sA[32 * FK_SEG * 12];  // DH matrices, then link transforms, of one segment
    const uint32_t p0 = blockIdx.x * 32, pl = tid & 31;
    if (blockIdx.x == 0)
      for (uint32_t i = tid; i < kStatSlots * 2; i += 256) stats[(i >> 1) * kStatStride + (i & 1)] = 0ull;
    float t[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) t[k] = base[k];
    for (uint32_t l0 = 0; l0 < nl; l0 += FK_SEG) {
      const uint32_t ns = nl - l0 < (uint32_t)FK_SEG ? nl - l0 : (uint32_t)FK_SEG;
      for (uint32_t i = tid; i < 32 * ns; i += 256) {
        const uint32_t pi = i & 31, ls = i >> 5, p = p0 + pi;
        if (p < n_poses) dh_matrix(dh + 5 * (l0 + ls), q[(uint64_t)p * nl + l0 + ls], sA + (pi * FK_SEG + ls) * 12);
      }
      __syncthreads();
      if (tid < 32 && p0 + tid < n_poses) {
        const uint64_t p = p0 + tid;
        for (uint32_t ls = 0; ls < ns; ++ls) {
          float* a = sA + (tid * FK_SEG + ls) * 12;
          float4* o = reinterpret_cast<float4*>(T + (p * nl + l0 + ls) * 12);
          o[0] = make_float4(t[0], t[1], t[2], t[3]);
          o[1] = make_float4(t[4], t[5], t[6], t[7]);
          o[2] = make_float4(t[8], t[9], t[10], t[11]);
          float nt[12];
          affine_mul(t, a, nt);
#pragma unroll
          for (int k = 0; k < 12; ++k) {
            a[k] = t[k];  // the slot now holds link l0 + ls's transform
            t[k] = nt[k];
          }
        }
      }
      __syncthreads();
      // link-level broad phase: lanes 0-31 and 32-63 of a wave are two links of the 32 poses
      for (uint32_t i0 = 0; i0 < 32 * (uint32_t)FK_SEG; i0 += 256) {
        const uint32_t i = i0 + tid, ls = i >> 5, p = p0 + pl;
        bool live = false;
        if (ls < ns && p < n_poses)
          live = box_may_hit(sA + (pl * FK_SEG + ls) * 12, lbox + 6 * (l0 + ls), side, inv_side, dx, dy, dz, sat,
                             bxn, byn);
        const uint64_t bal = wave_ballot(live);
        if ((tid & 31) == 0 && ls < ns)
          llive[(uint64_t)(l0 + ls) * lwords + blockIdx.x] = (uint32_t)(bal >> (tid & 32));
      }
      __syncthreads();
    }
    if (p0 + pl < n_poses && tid < 32) counts[p0 + pl] = 0u;
    for (uint32_t i = tid; i < 32 * 8; i += 256)
      if (p0 + i / 8 < n_poses) masks[(uint64_t)p0 * 8 + i] = 0u;
    return;
  }
  const uint32_t zb = blockIdx.x - fk_blocks, nzb = gridDim.x - fk_blocks;
  for (uint32_t sl = zb; sl < kTouchSlots; sl += nzb) {
    const uint32_t n = tcnt[sl * kTouchStride];
    const uint32_t* __restrict__ lst = tlist + (uint64_t)sl * tcap;
    for (uint32_t i = tid; i < n; i += 256) hits[((uint64_t)lst[i] << 8) | sl] = 0u;
    __syncthreads();  // every thread has read n
    if (tid == 0) tcnt[sl * kTouchStride] = 0u;
  }
}

// ---------------------------------------------------------------- main kernel
// block = (chunk c of C points, pose group g of PG = TB * PPT poses), TB threads.
// Broad phase: thread t takes poses g*PG + k*TB + t (k < PPT); a pose whose whole
// link missed every occupied brick (the prologue's mask) drops out, the others test
// the chunk's box; the block leaves when no pose is left, else the surviving poses
// are compacted (pose ids in LDS, their transforms in LDS windows of KW).  Narrow
// phase: thread t takes point t % C of the chunk and every (TB / C)-th surviving
// pose from slot t / C, U poses per round so that their brick loads overlap.
constexpr int KW = 256;  // surviving poses whose transforms sit in LDS at a time

#define GVX_BATCH_ARGS                                                                                            \
  const float *__restrict__ pts, const uint32_t *__restrict__ chunks, const float *__restrict__ boxes,            \
      uint32_t n_chunks, const float *__restrict__ T, uint32_t nl, uint32_t n_poses,                              \
      const uint32_t *__restrict__ sat, const uint64_t *__restrict__ bricks, const uint32_t *__restrict__ wrank,  \
      uint32_t dx, uint32_t dy, uint32_t dz, uint32_t bxn, uint32_t byn, float side, float inv_side,             \
      uint32_t *__restrict__ hits, uint32_t words, uint32_t *__restrict__ tlist,                                   \
      uint32_t *__restrict__ tcnt, uint32_t tcap, uint32_t *__restrict__ counts, uint32_t *__restrict__ masks,      \
      unsigned long long *__restrict__ stats, const uint32_t *__restrict__ llive, uint32_t lwords
#define GVX_BATCH_PASS                                                                                             \
  pts, chunks, boxes, n_chunks, T, nl, n_poses, sat, bricks, wrank, dx, dy, dz, bxn, byn, side, inv_side, hits,   \
      words, tlist, tcnt, tcap, counts, masks, stats, llive, lwords

// DD: dense environments, where neighbouring lanes (consecutive Morton-ordered points of
// one pose) often hit one voxel: only the first lane of such a run sends the atomic
template <int C, int TB, int PPT, bool DD>
__device__ __forceinline__ void collide_item(uint32_t item, float* sT, uint32_t* sP, uint32_t* sN, GVX_BATCH_ARGS)
{
  static_assert(C % 64 == 0 && TB % C == 0, "a wave covers 64 points of one pose slot");
  constexpr int PG = TB * PPT, NW = TB / 64, L = TB / C, U = 4;
  const uint32_t c = item % n_chunks, g = item / n_chunks;
  const uint32_t link = chunks[3 * c], start = chunks[3 * c + 1], cnt = chunks[3 * c + 2];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  bool live[PPT];
  bool any = false;
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const uint32_t p = g * PG + k * TB + threadIdx.x;
    live[k] = p < n_poses && ((llive[(uint64_t)link * lwords + (p >> 5)] >> (p & 31)) & 1u);
    any = any || live[k];
  }
  if (!__syncthreads_or(any)) return;
  // the chunk's point for the narrow phase, loaded early
  const uint32_t pi = threadIdx.x % C, slot = threadIdx.x / C;
  const bool valid = pi < cnt;
  float px = 0.f, py = 0.f, pz = 0.f;
  if (valid) {
    const uint64_t i = (uint64_t)start + pi;
    px = pts[3 * i]; py = pts[3 * i + 1]; pz = pts[3 * i + 2];
  }
  // broad phase: all transform loads first, then the (branch-free) box tests
  float bx[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) bx[k] = boxes[6 * c + k];
  float m[PPT][12];
#pragma unroll
  for (int k = 0; k < PPT; ++k)
    if (live[k]) {
      const float4* t4 = reinterpret_cast<const float4*>(T + ((uint64_t)(g * PG + k * TB + threadIdx.x) * nl + link) * 12);
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const float4 v = t4[r];
        m[k][4 * r] = v.x; m[k][4 * r + 1] = v.y; m[k][4 * r + 2] = v.z; m[k][4 * r + 3] = v.w;
      }
    }
  uint64_t bal[PPT];
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    live[k] = live[k] && box_may_hit(m[k], bx, side, inv_side, dx, dy, dz, sat, bxn, byn);
    bal[k] = wave_ballot(live[k]);
    if (lane == 0) sN[k * NW + wave] = (uint32_t)__popcll(bal[k]);
  }
  __syncthreads();
  uint32_t total = 0, base[PPT];
#pragma unroll
  for (int k = 0; k < PPT; ++k)
    for (int w = 0; w < NW; ++w) {
      if (w == wave) base[k] = total;
      total += sN[k * NW + w];
    }
  if (total == 0) return;  // uniform over the block
  if (threadIdx.x == 0) {
    unsigned long long* st = stats + (item % kStatSlots) * kStatStride;
    atomicAdd(st, (unsigned long long)total);
    atomicAdd(st + 1, (unsigned long long)total * cnt);
  }
#pragma unroll
  for (int k = 0; k < PPT; ++k)
    if (live[k]) {
      const uint32_t j = base[k] + (uint32_t)__popcll(bal[k] & ((1ull << lane) - 1ull));
      sP[j] = g * PG + k * TB + threadIdx.x;
      if (j < (uint32_t)KW) {
#pragma unroll
        for (int r = 0; r < 12; ++r) sT[12 * j + r] = m[k][r];
      }
    }
  __syncthreads();
  // narrow phase, window by window of KW surviving poses
  const uint32_t mbit = 1u << ((M_SV_START + link) & 31);
  const uint32_t mword = (M_SV_START + link) >> 5;
  for (uint32_t w0 = 0; w0 < total; w0 += KW) {
    if (w0 > 0) {  // transforms of a later window from the prologue's array (L2)
      __syncthreads();
      const uint32_t nwin = total - w0 < (uint32_t)KW ? total - w0 : (uint32_t)KW;
      for (uint32_t i = threadIdx.x; i < nwin * 12; i += TB)
        sT[i] = T[((uint64_t)sP[w0 + i / 12] * nl + link) * 12 + i % 12];
      __syncthreads();
    }
    const uint32_t wend = total - w0 < (uint32_t)KW ? total : w0 + KW;
    for (uint32_t j0 = w0 + slot; j0 < wend; j0 += L * U) {
      // a round: U poses' bitmap words and ranks loaded together, then their hit-matrix
      // atomics issued together, then the counts
      uint64_t w[U], widx[U];
      uint32_t rk[U], bit[U], old[U];
      bool in[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint32_t j = j0 + u * L;
        const uint32_t jj = j < wend ? j : wend - 1;
        float wx, wy, wz;
        transform(sT + 12 * (jj - w0), px, py, pz, wx, wy, wz);
        const uint32_t x = quantize(wx, side), y = quantize(wy, side), z = quantize(wz, side);
        in[u] = valid && j < wend && x < dx && y < dy && z < dz;
        const uint32_t br = in[u] ? ((z >> 3) * byn + (y >> 3)) * bxn + (x >> 3) : 0u;  // map bricks: < 2^32
        const uint64_t wi = (uint64_t)br * 8 + (in[u] ? (z & 7) : 0u);
        bit[u] = ((y & 7) << 3) | (x & 7);
        w[u] = bricks[wi];
        rk[u] = wrank[wi];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint32_t j = j0 + u * L;
        old[u] = ~0u;
        widx[u] = 0;
        const bool hit = in[u] && ((w[u] >> bit[u]) & 1ull);
        const uint32_t rank = hit ? rk[u] + (uint32_t)__popcll(w[u] & ((1ull << bit[u]) - 1ull)) : ~0u;
        bool lead = hit;
        if (DD) {  // a lane whose voxel is the previous lane's leaves the atomic (and the count) to it
          const uint32_t prev = __shfl_up(rank, 1, 64);
          lead = hit && !(lane > 0 && prev == rank);
        }
        if (lead) {
          const uint32_t pose = sP[j];
          widx[u] = (uint64_t)rank * words + (pose >> 5);
          old[u] = atomicOr(hits + widx[u], 1u << (pose & 31));
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint32_t j = j0 + u * L;
        if (j >= wend) break;  // uniform over the wave: a wave is one slot
        const uint32_t pose = sP[j];
        const bool hit = in[u] && ((w[u] >> bit[u]) & 1ull);
        const bool isnew = hit && ((old[u] >> (pose & 31)) & 1u) == 0u;  // old = ~0: no atomic of its own
        const uint64_t nb = wave_ballot(isnew);
        const uint64_t hb = wave_ballot(hit);
        if (lane == 0) {
          if (nb) atomicAdd(counts + pose, (uint32_t)__popcll(nb));
          if (hb) atomicOr(masks + (uint64_t)pose * 8 + mword, mbit);
        }
        // a word this call made non-zero goes to its slot's touched list
        if (hit && old[u] == 0u) {
          const uint32_t sl = (uint32_t)widx[u] & (kTouchSlots - 1u);
          const uint32_t at = atomicAdd(tcnt + sl * kTouchStride, 1u);
          tlist[(uint64_t)sl * tcap + at] = (uint32_t)(widx[u] >> 8);
        }
      }
    }
  }
}

// one work item per block, or (PERSIST) a grid of a few blocks per CU striding over
// the items
template <int C, int TB, int PPT, bool PERSIST, bool DD>
__global__ void __launch_bounds__(TB) k_batch_collide(uint32_t n_items, GVX_BATCH_ARGS)
{
  __shared__ float sT[KW * 12];
  __shared__ uint32_t sP[TB * PPT];
  __shared__ uint32_t sN[PPT * (TB / 64)];
  if (PERSIST) {
    for (uint32_t it = blockIdx.x; it < n_items; it += gridDim.x) {
      collide_item<C, TB, PPT, DD>(it, sT, sP, sN, GVX_BATCH_PASS);
      __syncthreads();
    }
  } else {
    collide_item<C, TB, PPT, DD>(blockIdx.x, sT, sP, sN, GVX_BATCH_PASS);
  }
}

// Order n link-frame points along a Morton curve of their bounding box (10 bits per
// axis), so that consecutive points form compact patches; non-finite points go last.
// The order of a link's points changes no result: counts are of distinct voxels and
// masks are per link.
static void morton_order(const float* p, uint32_t n, float* out)
{
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  auto finite = [&](uint32_t i) {
    return std::isfinite(p[3 * i]) && std::isfinite(p[3 * i + 1]) && std::isfinite(p[3 * i + 2]);
  };
  for (uint32_t i = 0; i < n; ++i)
    if (finite(i))
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], p[3 * i + a]);
        hi[a] = std::max(hi[a], p[3 * i + a]);
      }
  std::vector<std::pair<uint32_t, uint32_t>> key(n);
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t code = 0xFFFFFFFFu;
    if (finite(i)) {
      code = 0;
      for (int a = 0; a < 3; ++a) {
        const double ext = (double)hi[a] - (double)lo[a];
        const double u = ext > 0 ? ((double)p[3 * i + a] - (double)lo[a]) / ext * 1023.0 : 0.0;
        const uint32_t qa = (uint32_t)std::min(std::max(u, 0.0), 1023.0);
        for (int bit = 0; bit < 10; ++bit) code |= ((qa >> bit) & 1u) << (3 * bit + a);
      }
    }
    key[i] = {code, i};
  }
  std::sort(key.begin(), key.end());
  for (uint32_t k = 0; k < n; ++k)
    for (int a = 0; a < 3; ++a) out[3 * (uint64_t)k + a] = p[3 * (uint64_t)key[k].second + a];
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
  {
    const uint32_t ex = b->bxn + 1, ey = b->byn + 1, ez = b->bzn + 1;
    const uint64_t ns = (uint64_t)ex * ey * ez;
    hipLaunchKernelGGL(k_sat_init, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, s, b->bcount, b->bxn, b->byn,
                       b->bzn, b->bsat);
    const uint64_t lines[3] = {(uint64_t)ey * ez, (uint64_t)ex * ez, (uint64_t)ex * ey};
    for (int axis = 0; axis < 3; ++axis)
      hipLaunchKernelGGL(k_sat_scan, dim3((unsigned)((lines[axis] + 255) / 256)), dim3(256), 0, s, b->bsat, ex, ey,
                         ez, axis);
    GVX_HIP(hipGetLastError());
  }
  GVX_CHECK(scan_u32(b->bcount, b->bprefix, b->nbricks, b->bprefix + b->nbricks));
  hipLaunchKernelGGL(k_word_rank, dim3((unsigned)((b->nbricks + 255) / 256)), dim3(256), 0, s, b->bricks, b->bprefix,
                     b->nbricks, b->wrank);
  GVX_HIP(hipGetLastError());
  GVX_HIP(hipMemcpyAsync(rt().host_small, b->bprefix + b->nbricks, 4, hipMemcpyDeviceToHost, s));
  GVX_CHECK(sync());
  b->n_occ = *(uint32_t*)rt().host_small;
  const uint64_t words = (b->max_poses + 31) / 32;
  const uint64_t need = (uint64_t)(b->n_occ ? b->n_occ : 1) * words;
  // captured graphs hold the hit-matrix size of the previous environment
  for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
  b->graphs.clear();
  if (need > b->hits_words || !b->tlist) {
    GVX_CHECK(sync());
    dfree(b->hits);
    dfree(b->tlist);
    b->hits = nullptr;
    b->tlist = nullptr;
    b->hits_words = 0;
    GVX_CHECK(dalloc((void**)&b->hits, need * 4));
    b->hits_words = need;
    // a slot holds every word index of its residue: it cannot overflow
    const uint64_t cap = (need + kTouchSlots - 1) / kTouchSlots;
    if (cap > 0xFFFFFFFFull || need > (1ull << 40)) return fail(GVX_ERR_UNSUPPORTED, "hit matrix too large");
    b->tcap = (uint32_t)cap;
    GVX_CHECK(dalloc((void**)&b->tlist, (uint64_t)kTouchSlots * cap * 4));
    GVX_CHECK(fill_u32(b->hits, 0u, need));  // zero once; later calls clear only what they touched
    GVX_CHECK(fill_u32(b->tcnt, 0u, kTouchSlots * kTouchStride));
  }
  return GVX_OK;
}

// chunks of `block` consecutive points of one link (the last one of a link may be
// shorter), with their link-frame bounding boxes (NaN when a point of the chunk is
// not finite, which keeps every pose)
static int chunk_table(BatchS* b, int block, const ChunkTab** out)
{
  auto it = b->chunks.find(block);
  if (it != b->chunks.end()) {
    *out = &it->second;
    return GVX_OK;
  }
  std::vector<uint32_t> t;
  std::vector<float> box;
  uint64_t off = 0;
  for (uint32_t l = 0; l < b->n_links; ++l) {
    for (uint32_t s = 0; s < b->sizes[l]; s += block) {
      const uint32_t cnt = b->sizes[l] - s < (uint32_t)block ? b->sizes[l] - s : (uint32_t)block;
      t.push_back(l);
      t.push_back((uint32_t)(off + s));
      t.push_back(cnt);
      float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
      bool finite = true;
      for (uint32_t i = 0; i < cnt; ++i) {
        const float* p = &b->hpts[3 * (off + s + i)];
        for (int a = 0; a < 3; ++a) {
          finite = finite && std::isfinite(p[a]);
          lo[a] = std::min(lo[a], p[a]);
          hi[a] = std::max(hi[a], p[a]);
        }
      }
      for (int a = 0; a < 3; ++a) box.push_back(finite ? lo[a] : NAN);
      for (int a = 0; a < 3; ++a) box.push_back(finite ? hi[a] : NAN);
    }
    off += b->sizes[l];
  }
  ChunkTab ct;
  ct.n = (uint32_t)(t.size() / 3);
  GVX_CHECK(dalloc((void**)&ct.tab, t.size() * 4 + 4));
  if (dalloc((void**)&ct.box, box.size() * 4 + 4) != GVX_OK) {
    dfree(ct.tab);
    return GVX_ERR_OOM;
  }
  if (!t.empty()) {
    GVX_HIP(hipMemcpyAsync(ct.tab, t.data(), t.size() * 4, hipMemcpyHostToDevice, rt().stream));
    GVX_HIP(hipMemcpyAsync(ct.box, box.data(), box.size() * 4, hipMemcpyHostToDevice, rt().stream));
  }
  GVX_CHECK(sync());
  *out = &(b->chunks[block] = ct);
  return GVX_OK;
}

// enqueue one batched evaluation of n poses: q (n x n_links) in, counts (n) and
// masks (n x 8 words) out, all device memory
static int batch_enqueue(BatchS* b, uint32_t n, int sched, const float* q, uint32_t* counts, uint32_t* masks)
{
  hipStream_t s = rt().stream;
  const Schedule& S = kSchedules[sched];
  const ChunkTab* ct = nullptr;
  GVX_CHECK(chunk_table(b, S.chunk, &ct));
  const uint32_t words = (b->max_poses + 31) / 32;
  const bool narrow = ct->n != 0 && b->n_occ != 0;
  const float inv_side = 1.0f / b->side;
  const uint32_t fk_blocks = (n + 31) / 32;
  // clearing blocks: one per touched-word slot
  const uint64_t zb = kTouchSlots;
  hipLaunchKernelGGL(k_batch_prologue, dim3((unsigned)(fk_blocks + zb)), dim3(256), 0, s, b->hits,
                     b->tlist, b->tcnt, b->tcap, counts, masks, b->dh, b->base, q, n, b->n_links, b->T, b->stats,
                     b->lbox, b->llive, b->lwords, fk_blocks, b->bsat, b->dims[0], b->dims[1], b->dims[2], b->bxn,
                     b->byn, b->side, inv_side);
  GVX_HIP(hipGetLastError());
  b->last_tests = narrow ? (uint64_t)ct->n * n : 0;
  if (!narrow) return GVX_OK;
  const uint64_t pgs = (uint64_t)S.block * S.ppt;
  const uint64_t grid = (uint64_t)ct->n * ((n + pgs - 1) / pgs);
  if (grid > 0x7FFFFFFFull) return fail(GVX_ERR_UNSUPPORTED, "too many (chunk, pose group) work items");
  uint64_t lgrid = grid;
  if (S.persist) {
    const uint64_t pgrid = (uint64_t)(rt().num_cus > 0 ? rt().num_cus : 256) * S.persist;
    lgrid = grid < pgrid ? grid : pgrid;
  }
  // environments with more than 1/64 of their voxels occupied take the run-merging kernels
  const bool dd = (uint64_t)b->n_occ * 64 > (uint64_t)b->dims[0] * b->dims[1] * b->dims[2];
#define LAUNCH_DD(CS, BS, PP, PE, DD)                                                                            \
  hipLaunchKernelGGL((k_batch_collide<CS, BS, PP, PE, DD>), dim3((unsigned)lgrid), dim3(BS), 0, s, (uint32_t)grid, \
                     b->pts, ct->tab, ct->box, ct->n, b->T, b->n_links, n, b->bsat, b->bricks, b->wrank, b->dims[0], \
                     b->dims[1], b->dims[2], b->bxn, b->byn, b->side, inv_side, b->hits, words, b->tlist, b->tcnt,  \
                     b->tcap, counts, masks, b->stats, b->llive, b->lwords)
#define LAUNCH(CS, BS, PP, PE)                                                                                   \
  if (dd) LAUNCH_DD(CS, BS, PP, PE, true); else LAUNCH_DD(CS, BS, PP, PE, false)
  switch (sched) {
    case 0: LAUNCH(128, 512, 1, false); break;
    case 1: LAUNCH(128, 256, 1, false); break;
    case 2: LAUNCH(64, 256, 1, false); break;
    case 3: LAUNCH(256, 256, 1, false); break;
    case 4: LAUNCH(128, 512, 1, true); break;
    case 5: LAUNCH(128, 256, 1, true); break;
    case 6: LAUNCH(64, 256, 1, true); break;
    default: LAUNCH(128, 256, 2, false); break;
  }
#undef LAUNCH
#undef LAUNCH_DD
  GVX_HIP(hipGetLastError());
  return GVX_OK;
}

// one batched call on device buffers: replay of the graph captured for these
// buffers (captured on first use; at most 64 kept), or direct launches
static int batch_run(BatchS* b, uint32_t n, int sched, const float* q, uint32_t* counts, uint32_t* masks)
{
  hipStream_t s = rt().stream;
  const int ks = sched % kKernelSchedules;
  const ChunkTab* ct;
  GVX_CHECK(chunk_table(b, kSchedules[ks].chunk, &ct));  // built before any capture
  if (sched < kKernelSchedules) {
    ProfScope ps("batch_collide");
    return batch_enqueue(b, n, ks, q, counts, masks);
  }
  auto key = std::make_tuple(n, ks, (const void*)q, (void*)counts, (void*)masks);
  auto it = b->graphs.find(key);
  if (it == b->graphs.end()) {
    if (b->graphs.size() >= 64) {
      GVX_CHECK(sync());
      for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
      b->graphs.clear();
    }
    hipGraph_t graph;
    GVX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int st = batch_enqueue(b, n, ks, q, counts, masks);
    hipError_t e = hipStreamEndCapture(s, &graph);
    if (st != GVX_OK) return st;
    if (e != hipSuccess) return fail(GVX_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    hipGraphExec_t exec;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(GVX_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    it = b->graphs.emplace(key, exec).first;
  }
  b->last_tests = (ct->n != 0 && b->n_occ != 0) ? (uint64_t)ct->n * n : 0;
  ProfScope ps("batch_collide");
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
  if ((s = dalloc((void**)&b->wrank, b->nbricks * 32)) != GVX_OK) return bail(s);
  if ((s = dalloc((void**)&b->bsat, (uint64_t)(b->bxn + 1) * (b->byn + 1) * (b->bzn + 1) * 4)) != GVX_OK)
    return bail(s);
  if ((s = dalloc((void**)&b->pts, tot * 12 + 16)) != GVX_OK) return bail(s);
  if ((s = dalloc((void**)&b->dh, n_links * 20)) != GVX_OK) return bail(s);
  if ((s = dalloc((void**)&b->base, 48)) != GVX_OK) return bail(s);
  if ((s = dalloc((void**)&b->q, (uint64_t)max_poses * n_links * 4)) != GVX_OK) return bail(s);
  if ((s = dalloc((void**)&b->T, (uint64_t)max_poses * n_links * 48)) != GVX_OK) return bail(s);
  if ((s = dalloc((void**)&b->counts, (uint64_t)max_poses * 4)) != GVX_OK) return bail(s);
  if ((s = dalloc((void**)&b->masks, (uint64_t)max_poses * 32)) != GVX_OK) return bail(s);
  if ((s = dalloc((void**)&b->stats, kStatSlots * kStatStride * 8)) != GVX_OK) return bail(s);
  if ((s = dalloc((void**)&b->tcnt, kTouchSlots * kTouchStride * 4)) != GVX_OK) return bail(s);
  b->lwords = (max_poses + 31) / 32;
  if ((s = dalloc((void**)&b->llive, (uint64_t)n_links * b->lwords * 4)) != GVX_OK) return bail(s);
  if ((s = dalloc((void**)&b->lbox, (uint64_t)n_links * 24)) != GVX_OK) return bail(s);
  float ident[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
  hipStream_t st = rt().stream;
  b->hpts.resize(tot * 3);
  for (uint32_t l = 0, off = 0; l < n_links; off += link_sizes[l], ++l)
    morton_order(link_pts + 3 * (uint64_t)off, link_sizes[l], b->hpts.data() + 3 * (uint64_t)off);
  std::vector<float> lb(6 * (size_t)n_links);
  for (uint32_t l = 0, off = 0; l < n_links; off += link_sizes[l], ++l) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    bool finite = true;
    for (uint32_t i = 0; i < link_sizes[l]; ++i)
      for (int a = 0; a < 3; ++a) {
        const float v = b->hpts[3 * ((uint64_t)off + i) + a];
        finite = finite && std::isfinite(v);
        lo[a] = std::min(lo[a], v);
        hi[a] = std::max(hi[a], v);
      }
    for (int a = 0; a < 3; ++a) {
      lb[6 * l + a] = finite ? lo[a] : NAN;  // an empty link keeps +-inf: its magnitude test keeps every pose
      lb[6 * l + 3 + a] = finite ? hi[a] : NAN;
    }
  }
  if (hipMemcpyAsync(b->lbox, lb.data(), lb.size() * 4, hipMemcpyHostToDevice, st) != hipSuccess)
    return bail(fail(GVX_ERR_HIP, "copy link boxes"));
  if (tot && hipMemcpyAsync(b->pts, b->hpts.data(), tot * 12, hipMemcpyHostToDevice, st) != hipSuccess)
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
  for (auto& kv : b->chunks) {
    dfree(kv.second.tab);
    dfree(kv.second.box);
  }
  void* bufs[] = {b->bricks, b->bcount, b->bprefix, b->wrank, b->bsat, b->pts, b->dh, b->base,
                  b->q,      b->T,      b->hits,    b->counts, b->masks, b->stats, b->llive, b->lbox,
                  b->tlist,  b->tcnt};
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
                              int out_mem, int sched)
{
  if (!b) return fail(GVX_ERR_BAD_ARG, "null batch");
  if (n > b->max_poses) return fail(GVX_ERR_BAD_ARG, "n_poses > max_poses");
  if (n && !q) return fail(GVX_ERR_BAD_ARG, "null q");
  if (n == 0) return GVX_OK;
  std::lock_guard<std::recursive_mutex> g0(b->mu);
  std::lock_guard<std::recursive_mutex> g(b->env->mu);
  std::lock_guard<std::recursive_mutex> g2(rt().mu);
  if (mem == GVX_MEM_DEVICE && out_mem == GVX_MEM_DEVICE && counts && masks && ((uintptr_t)q & 3) == 0 &&
      ((uintptr_t)counts & 3) == 0 && ((uintptr_t)masks & 3) == 0) {
    // everything in device memory: the two kernels read and write the caller's buffers
    GVX_CHECK(batch_run(b, n, sched, q, counts, reinterpret_cast<uint32_t*>(masks)));
    return sync();
  }
  GVX_CHECK(copy_in(b->q, q, (uint64_t)n * b->n_links * 4, mem));
  GVX_CHECK(batch_run(b, n, sched, b->q, b->counts, b->masks));
  if (counts) GVX_CHECK(copy_out(counts, b->counts, (uint64_t)n * 4, out_mem));
  if (masks) GVX_CHECK(copy_out(masks, b->masks, (uint64_t)n * 32, out_mem));
  return sync();
}

int gvx_batch_collide(gvx_batch b, const float* q, uint32_t n, int mem, uint32_t* counts, uint8_t* masks, int out_mem)
{
  if (!b) return fail(GVX_ERR_BAD_ARG, "null batch");
  return batch_collide_impl(b, q, n, mem, counts, masks, out_mem, b->schedule);
}

int gvx_batch_collide_async(gvx_batch b, const float* q_dev, uint32_t n, uint32_t* counts_dev, uint8_t* masks_dev)
{
  if (!b) return fail(GVX_ERR_BAD_ARG, "null batch");
  if (n > b->max_poses) return fail(GVX_ERR_BAD_ARG, "n_poses > max_poses");
  if (n == 0) return GVX_OK;
  if (!q_dev || !counts_dev || !masks_dev) return fail(GVX_ERR_BAD_ARG, "null device buffer");
  if (((uintptr_t)q_dev & 3) || ((uintptr_t)counts_dev & 3) || ((uintptr_t)masks_dev & 3))
    return fail(GVX_ERR_BAD_ARG, "device buffers must be 4-byte aligned");
  std::lock_guard<std::recursive_mutex> g0(b->mu);
  std::lock_guard<std::recursive_mutex> g(b->env->mu);
  std::lock_guard<std::recursive_mutex> g2(rt().mu);
  return batch_run(b, n, b->schedule, q_dev, counts_dev, reinterpret_cast<uint32_t*>(masks_dev));
}

int gvx_batch_last_stats(gvx_batch b, uint64_t* tests, uint64_t* live_pairs, uint64_t* live_point_poses)
{
  if (!b) return fail(GVX_ERR_BAD_ARG, "null batch");
  std::lock_guard<std::recursive_mutex> g0(b->mu);
  std::lock_guard<std::recursive_mutex> g2(rt().mu);
  std::vector<unsigned long long> h(kStatSlots * kStatStride, 0ull);
  if (b->last_tests) {
    GVX_HIP(hipMemcpyAsync(h.data(), b->stats, h.size() * 8, hipMemcpyDeviceToHost, rt().stream));
    GVX_CHECK(sync());
  }
  uint64_t sum[2] = {0, 0};
  for (uint32_t k = 0; k < kStatSlots; ++k) {
    sum[0] += h[k * kStatStride];
    sum[1] += h[k * kStatStride + 1];
  }
  if (tests) *tests = b->last_tests;
  if (live_pairs) *live_pairs = sum[0];
  if (live_point_poses) *live_point_poses = sum[1];
  return GVX_OK;
}

int gvx_sync(void)
{
  GVX_CHECK(ensure_init());
  return sync();
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
  // tune on the poses in b->q: those of the last call that copied its joint values in
  // (a host-memory q), zeros before any
  hipEvent_t e0, e1;
  GVX_HIP(hipEventCreate(&e0));
  GVX_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  int best_s = b->schedule;
  for (int s = 0; s < kNumSchedules; ++s) {
    GVX_CHECK(batch_run(b, n, s, b->q, b->counts, b->masks));  // warm-up (and capture)
    GVX_HIP(hipEventRecord(e0, rt().stream));
    for (int r = 0; r < 10; ++r) GVX_CHECK(batch_run(b, n, s, b->q, b->counts, b->masks));
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
