// gvx_lists.hip -- BitVectorVoxelList: insert, LSD radix sort, run-length OR / unique,
// merge, list x map and list x list collision.
#include "gvx_internal.hpp"

namespace gvx {

struct ListS {
  uint32_t dims[3];
  float side;
  uint64_t n = 0, cap = 0;
  uint32_t* ids = nullptr;
  uint32_t* coords = nullptr;
  uint32_t* bits = nullptr;  // 8 words per entry
  // scratch, grown on demand
  uint64_t tcap = 0;
  uint32_t *k2 = nullptr, *v1 = nullptr, *v2 = nullptr, *flags = nullptr, *pos = nullptr;
  uint32_t* hist = nullptr;
  uint64_t hcap = 0;
  uint32_t *ids_t = nullptr, *coords_t = nullptr, *bits_t = nullptr;
  // every id < dx*dy*dz <= 2^32 and every coordinate is its id's: duplicates can then
  // be merged through a dense occupancy bitmap instead of a sort
  bool ids_in_range = true;
  std::recursive_mutex mu;
};

}  // namespace gvx

struct gvx_list_s : gvx::ListS {};

namespace gvx {

static uint64_t nvox3(const uint32_t d[3]) { return (uint64_t)d[0] * d[1] * d[2]; }

static int list_reserve(ListS* L, uint64_t want)
{
  if (want <= L->cap) return GVX_OK;
  uint64_t nc = L->cap ? L->cap : 1024;
  while (nc < want) nc *= 2;
  uint32_t *ni = nullptr, *nco = nullptr, *nb = nullptr;
  GVX_CHECK(dalloc((void**)&ni, nc * 4));
  GVX_CHECK(dalloc((void**)&nco, nc * 12));
  GVX_CHECK(dalloc((void**)&nb, nc * 32));
  if (L->n) {
    GVX_HIP(hipMemcpyAsync(ni, L->ids, L->n * 4, hipMemcpyDeviceToDevice, rt().stream));
    GVX_HIP(hipMemcpyAsync(nco, L->coords, L->n * 12, hipMemcpyDeviceToDevice, rt().stream));
    GVX_HIP(hipMemcpyAsync(nb, L->bits, L->n * 32, hipMemcpyDeviceToDevice, rt().stream));
  }
  GVX_CHECK(sync());
  dfree(L->ids); dfree(L->coords); dfree(L->bits);
  L->ids = ni; L->coords = nco; L->bits = nb;
  L->cap = nc;
  return GVX_OK;
}

static int tmp_reserve(ListS* L, uint64_t want)
{
  if (want <= L->tcap) return GVX_OK;
  GVX_CHECK(sync());
  uint32_t** bufs[] = {&L->k2, &L->v1, &L->v2, &L->flags, &L->pos, &L->ids_t, &L->coords_t, &L->bits_t};
  for (auto b : bufs) { dfree(*b); *b = nullptr; }
  uint64_t nc = want < 1024 ? 1024 : want + want / 4;
  GVX_CHECK(dalloc((void**)&L->k2, nc * 4));
  GVX_CHECK(dalloc((void**)&L->v1, nc * 4));
  GVX_CHECK(dalloc((void**)&L->v2, nc * 4));
  GVX_CHECK(dalloc((void**)&L->flags, (nc + 1) * 4));
  GVX_CHECK(dalloc((void**)&L->pos, (nc + 1) * 4));
  GVX_CHECK(dalloc((void**)&L->ids_t, nc * 4));
  GVX_CHECK(dalloc((void**)&L->coords_t, nc * 12));
  GVX_CHECK(dalloc((void**)&L->bits_t, nc * 32));
  L->tcap = nc;
  return GVX_OK;
}

// ---------------------------------------------------------------- insert kernels
// kernelInsertGlobalPointCloud / kernelInsertMetaPointCloud (VoxelListOperations.hpp:41-140):
// (id, coord, voxel) per point; keep flag = inside the reference map (remove_out_of_bounds,
// TemplateVoxelList.hpp:297-341).
__global__ void __launch_bounds__(256) k_list_points(const float* __restrict__ pts, uint64_t n, uint32_t dx, uint32_t dy,
                                                     uint32_t dz, float side, const uint32_t* __restrict__ ends,
                                                     uint32_t ncl, const uint32_t* __restrict__ meanings,
                                                     uint32_t meaning, uint32_t* __restrict__ ids,
                                                     uint32_t* __restrict__ coords, uint32_t* __restrict__ bits,
                                                     uint32_t* __restrict__ keep)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    uint32_t m = meaning;
    if (ends) {
      uint32_t lo = 0, hi = ncl - 1;
      while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if ((uint64_t)ends[mid] > i) hi = mid; else lo = mid + 1;
      }
      m = meanings[lo];
    }
    const uint32_t x = quantize(pts[3 * i], side), y = quantize(pts[3 * i + 1], side),
                   z = quantize(pts[3 * i + 2], side);
    coords[3 * i] = x; coords[3 * i + 1] = y; coords[3 * i + 2] = z;
    ids[i] = z * dx * dy + y * dx + x;  // uint32 arithmetic as getVoxelIndexUnsigned
    uint4* b = reinterpret_cast<uint4*>(bits + 8 * i);
    uint32_t w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    w[m >> 5] = 1u << (m & 31);
    b[0] = make_uint4(w[0], w[1], w[2], w[3]);
    b[1] = make_uint4(w[4], w[5], w[6], w[7]);
    keep[i] = (x < dx && y < dy && z < dz) ? 1u : 0u;
  }
}

__global__ void __launch_bounds__(256) k_list_coords(const uint32_t* __restrict__ c, uint64_t n, uint32_t dx, uint32_t dy,
                                                     uint32_t dz, uint32_t meaning, uint32_t* __restrict__ ids,
                                                     uint32_t* __restrict__ coords, uint32_t* __restrict__ bits,
                                                     uint32_t* __restrict__ keep)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t x = c[3 * i], y = c[3 * i + 1], z = c[3 * i + 2];
    coords[3 * i] = x; coords[3 * i + 1] = y; coords[3 * i + 2] = z;
    ids[i] = z * dx * dy + y * dx + x;
    uint4* b = reinterpret_cast<uint4*>(bits + 8 * i);
    uint32_t w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    w[meaning >> 5] = 1u << (meaning & 31);
    b[0] = make_uint4(w[0], w[1], w[2], w[3]);
    b[1] = make_uint4(w[4], w[5], w[6], w[7]);
    keep[i] = (x < dx && y < dy && z < dz) ? 1u : 0u;
  }
}

// stable compaction of entries by keep flags (pos = exclusive scan of keep)
__global__ void __launch_bounds__(256) k_compact(const uint32_t* __restrict__ ids, const uint32_t* __restrict__ coords,
                                                 const uint32_t* __restrict__ bits, const uint32_t* __restrict__ keep,
                                                 const uint32_t* __restrict__ pos, uint64_t n,
                                                 uint32_t* __restrict__ oids, uint32_t* __restrict__ ocoords,
                                                 uint32_t* __restrict__ obits)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if (!keep[i]) continue;
    const uint64_t o = pos[i];
    oids[o] = ids[i];
    ocoords[3 * o] = coords[3 * i]; ocoords[3 * o + 1] = coords[3 * i + 1]; ocoords[3 * o + 2] = coords[3 * i + 2];
    const uint4* b = reinterpret_cast<const uint4*>(bits + 8 * i);
    uint4* ob = reinterpret_cast<uint4*>(obits + 8 * o);
    ob[0] = b[0];
    ob[1] = b[1];
  }
}

// entries [0, *count) of (ids, coords, bits) -> (oids, ocoords, obits); the count is
// read on the device, so the host need not know it before the copy (cap >= *count)
__global__ void __launch_bounds__(256) k_copy_entries(const uint32_t* __restrict__ count, uint64_t cap,
                                                      const uint32_t* __restrict__ ids,
                                                      const uint32_t* __restrict__ coords,
                                                      const uint32_t* __restrict__ bits, uint32_t* __restrict__ oids,
                                                      uint32_t* __restrict__ ocoords, uint32_t* __restrict__ obits)
{
  const uint64_t n = *count < cap ? *count : cap;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    oids[i] = ids[i];
    ocoords[3 * i] = coords[3 * i]; ocoords[3 * i + 1] = coords[3 * i + 1]; ocoords[3 * i + 2] = coords[3 * i + 2];
    const uint4* b = reinterpret_cast<const uint4*>(bits + 8 * i);
    uint4* ob = reinterpret_cast<uint4*>(obits + 8 * i);
    ob[0] = b[0];
    ob[1] = b[1];
  }
}

// ---------------------------------------------------------------- LSD radix sort (8-bit digits)
constexpr int RS_BLOCK = 256, RS_ROUNDS = 16, RS_TILE = RS_BLOCK * RS_ROUNDS;

__global__ void k_iota(uint32_t* __restrict__ v, uint64_t n)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) v[i] = (uint32_t)i;
}

// one atomic per block (a shared counter takes one atomic at a time)
__global__ void __launch_bounds__(256) k_max_u32(const uint32_t* __restrict__ k, uint64_t n, uint32_t* __restrict__ out)
{
  __shared__ uint32_t wm[4];
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint32_t m = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) m = k[i] > m ? k[i] : m;
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t y = __shfl_xor(m, o);
    m = y > m ? y : m;
  }
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) m = wm[w] > m ? wm[w] : m;
    atomicMax(out, m);
  }
}

// per-tile digit histogram, written digit-major: hist[d * tiles + tile]
__global__ void __launch_bounds__(RS_BLOCK) k_rs_hist(const uint32_t* __restrict__ keys, uint64_t n, int shift,
                                                      uint32_t* __restrict__ hist, uint32_t tiles)
{
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
  for (int r = 0; r < RS_ROUNDS; ++r) {
    uint64_t i = base + (uint64_t)r * RS_BLOCK + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xFF], 1u);
  }
  __syncthreads();
  hist[(uint64_t)threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

// stable scatter: within a tile, keys keep their input order per digit
__global__ void __launch_bounds__(RS_BLOCK) k_rs_scatter(const uint32_t* __restrict__ keys,
                                                         const uint32_t* __restrict__ vals, uint64_t n, int shift,
                                                         const uint32_t* __restrict__ offs, uint32_t tiles,
                                                         uint32_t* __restrict__ okeys, uint32_t* __restrict__ ovals)
{
  __shared__ uint32_t run[256];             // running count per digit in this tile
  __shared__ uint32_t wcnt[RS_BLOCK / 64][256];  // per-wave digit counts of the current round
  __shared__ uint32_t gbase[256];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  run[tid] = 0;
  gbase[tid] = offs[(uint64_t)tid * tiles + blockIdx.x];
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
  for (int r = 0; r < RS_ROUNDS; ++r) {
#pragma unroll
    for (int k = 0; k < RS_BLOCK / 64; ++k) wcnt[k][tid] = 0;
    __syncthreads();
    const uint64_t i = base + (uint64_t)r * RS_BLOCK + tid;
    const bool valid = i < n;
    uint32_t key = valid ? keys[i] : 0u;
    uint32_t d = (key >> shift) & 0xFF;
    uint64_t peers = wave_ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      uint64_t bb = wave_ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? bb : ~bb;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt);
    if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pre = 0;
      for (int k = 0; k < w; ++k) pre += wcnt[k][d];
      const uint32_t o = gbase[d] + run[d] + pre + rank;
      okeys[o] = key;
      ovals[o] = vals[i];
    }
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int k = 0; k < RS_BLOCK / 64; ++k) add += wcnt[k][tid];
    run[tid] += add;
    __syncthreads();
  }
}

// run heads after sorting
__global__ void k_heads(const uint32_t* __restrict__ keys, uint64_t n, uint32_t* __restrict__ head)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    head[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

// make_unique (TemplateVoxelList.hpp:143-212): first entry of each run keeps its
// coordinate and receives the OR (BitVoxel::reduce) of the run's payloads.
__global__ void __launch_bounds__(256) k_unique(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ perm,
                                                const uint32_t* __restrict__ head, const uint32_t* __restrict__ pos,
                                                uint64_t n, const uint32_t* __restrict__ coords,
                                                const uint32_t* __restrict__ bits, uint32_t* __restrict__ oids,
                                                uint32_t* __restrict__ ocoords, uint32_t* __restrict__ obits)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if (!head[i]) continue;
    const uint64_t o = pos[i];
    const uint32_t key = keys[i];
    const uint32_t p = perm[i];
    oids[o] = key;
    ocoords[3 * o] = coords[3 * (uint64_t)p];
    ocoords[3 * o + 1] = coords[3 * (uint64_t)p + 1];
    ocoords[3 * o + 2] = coords[3 * (uint64_t)p + 2];
    const uint4* b = reinterpret_cast<const uint4*>(bits + 8 * (uint64_t)p);
    uint4 a0 = b[0], a1 = b[1];
    for (uint64_t j = i + 1; j < n && keys[j] == key; ++j) {
      const uint4* c = reinterpret_cast<const uint4*>(bits + 8 * (uint64_t)perm[j]);
      uint4 c0 = c[0], c1 = c[1];
      a0.x |= c0.x; a0.y |= c0.y; a0.z |= c0.z; a0.w |= c0.w;
      a1.x |= c1.x; a1.y |= c1.y; a1.z |= c1.z; a1.w |= c1.w;
    }
    uint4* ob = reinterpret_cast<uint4*>(obits + 8 * o);
    ob[0] = a0;
    ob[1] = a1;
  }
}

static int read_u32(const uint32_t* dev, uint32_t* out)
{
  GVX_HIP(hipMemcpyAsync(rt().host_small, dev, 4, hipMemcpyDeviceToHost, rt().stream));
  GVX_CHECK(sync());
  *out = *(uint32_t*)rt().host_small;
  return GVX_OK;
}

// Stable LSD radix sort of (keys, vals) in place (n <= 2^32).
static int radix_sort(ListS* L, uint32_t* keys, uint32_t* vals, uint64_t n)
{
  if (n <= 1) return GVX_OK;
  hipStream_t s = rt().stream;
  uint32_t mx = 0;
  if (L->ids_in_range && nvox3(L->dims) <= (1ull << 32)) {
    mx = (uint32_t)(nvox3(L->dims) - 1);  // every key is an id below the map size
  } else {
    uint32_t* dmax = (uint32_t*)rt().dev_small + 16;
    GVX_CHECK(fill_u32(dmax, 0u, 4));
    hipLaunchKernelGGL(k_max_u32, dim3(grid_for(n, 256, 1024)), dim3(256), 0, s, keys, n, dmax);
    GVX_HIP(hipGetLastError());
    GVX_CHECK(read_u32(dmax, &mx));
  }
  int bits = 0;
  while (bits < 32 && (mx >> bits) != 0) bits += 8;
  const uint32_t tiles = (uint32_t)((n + RS_TILE - 1) / RS_TILE);
  uint64_t hn = (uint64_t)tiles * 256;
  if (L->hcap < 2 * hn) {
    GVX_CHECK(sync());
    dfree(L->hist);
    GVX_CHECK(dalloc((void**)&L->hist, 2 * hn * 4));
    L->hcap = 2 * hn;
  }
  uint32_t *ka = keys, *va = vals, *kb = L->k2, *vb = L->v2;
  for (int sh = 0; sh < bits; sh += 8) {
    hipLaunchKernelGGL(k_rs_hist, dim3(tiles), dim3(RS_BLOCK), 0, s, ka, n, sh, L->hist, tiles);
    GVX_HIP(hipGetLastError());
    GVX_CHECK(scan_u32(L->hist, L->hist + hn, hn, nullptr));
    hipLaunchKernelGGL(k_rs_scatter, dim3(tiles), dim3(RS_BLOCK), 0, s, ka, va, n, sh, L->hist + hn, tiles, kb, vb);
    GVX_HIP(hipGetLastError());
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  if (ka != keys) {
    GVX_HIP(hipMemcpyAsync(keys, ka, n * 4, hipMemcpyDeviceToDevice, s));
    GVX_HIP(hipMemcpyAsync(vals, va, n * 4, hipMemcpyDeviceToDevice, s));
  }
  return GVX_OK;
}

// ---------------------------------------------------------------- dense-bitmap unique
// For lists whose ids lie in [0, dx*dy*dz): mark every id in a bitmap of the map
// (one bit per voxel, zero at rest), rank = occupied voxels before the id (prefix of
// the words' popcounts + popcount inside the word), so entries land sorted by id
// with duplicates on one slot, where their payloads are ORed and their (identical)
// id and coordinates written.  Same output as sort + run-length OR.
__global__ void __launch_bounds__(256) k_lu_mark(const uint32_t* __restrict__ ids, const uint32_t* __restrict__ keep,
                                                 uint64_t keep_from, uint64_t n, unsigned long long* __restrict__ bm)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const bool in = i < keep_from || keep[i - keep_from] != 0u;
    const uint32_t id = in ? ids[i] : 0u;
    // neighbouring points of an image often share a voxel: one atomic per run of lanes
    if (wave_run_head(in, id, 0u)) atomicOr(bm + (id >> 6), 1ull << (id & 63));
  }
}

__global__ void __launch_bounds__(256) k_lu_popc(const unsigned long long* __restrict__ bm, uint64_t words,
                                                 uint32_t* __restrict__ pc)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < words; w += stride)
    pc[w] = (uint32_t)__popcll(bm[w]);
}

__global__ void __launch_bounds__(256) k_lu_scatter(const uint32_t* __restrict__ ids, const uint32_t* __restrict__ coords,
                                                    const uint32_t* __restrict__ bits, const uint32_t* __restrict__ keep,
                                                    uint64_t keep_from, uint64_t n,
                                                    const unsigned long long* __restrict__ bm,
                                                    const uint32_t* __restrict__ wpre, uint32_t* __restrict__ oids,
                                                    uint32_t* __restrict__ ocoords, uint32_t* __restrict__ obits)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if (i >= keep_from && !keep[i - keep_from]) continue;
    const uint32_t id = ids[i];
    const uint64_t o = wpre[id >> 6] + (uint32_t)__popcll(bm[id >> 6] & ((1ull << (id & 63)) - 1ull));
    oids[o] = id;
    ocoords[3 * o] = coords[3 * i];
    ocoords[3 * o + 1] = coords[3 * i + 1];
    ocoords[3 * o + 2] = coords[3 * i + 2];
    const uint4* b = reinterpret_cast<const uint4*>(bits + 8 * i);
    const uint4 a0 = b[0], a1 = b[1];
    const uint32_t v[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (v[k]) atomicOr(obits + 8 * o + k, v[k]);
  }
}

__global__ void __launch_bounds__(256) k_lu_clear(const uint32_t* __restrict__ ids, const uint32_t* __restrict__ keep,
                                                  uint64_t keep_from, uint64_t n, unsigned long long* __restrict__ bm)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    if (i < keep_from || keep[i - keep_from]) bm[ids[i] >> 6] = 0ull;  // dropped entries may hold any id
}

// the bitmap (zero at rest) and the per-word counts and prefix, shared by all lists
// (list calls hold the runtime lock)
static int lu_buffers(uint64_t words, unsigned long long** bm, uint32_t** pc, uint32_t** wpre)
{
  static unsigned long long* b = nullptr;
  static uint32_t *c = nullptr, *p = nullptr;
  static uint64_t cap = 0;
  if (cap < words) {
    GVX_CHECK(sync());
    dfree(b); dfree(c); dfree(p);
    b = nullptr; c = nullptr; p = nullptr;
    cap = 0;
    GVX_CHECK(dalloc((void**)&b, words * 8));
    GVX_CHECK(dalloc((void**)&c, words * 4));
    GVX_CHECK(dalloc((void**)&p, (words + 1) * 4));
    GVX_CHECK(fill_u32(b, 0u, words * 2));
    cap = words;
  }
  *bm = b; *pc = c; *wpre = p;
  return GVX_OK;
}

static bool lu_suits(const ListS* L, uint64_t entries)
{
  const uint64_t nv = nvox3(L->dims);
  if (!L->ids_in_range || nv > (1ull << 32) || entries == 0) return false;
  const uint64_t words = (nv + 63) / 64;
  return words <= (1ull << 16) || words <= 64 * entries;
}

// entries [0, n) of L (entries from keep_from on only where keep[i - keep_from]) ->
// the sorted, duplicate-free list
static int lu_unique(ListS* L, uint64_t n, const uint32_t* keep, uint64_t keep_from)
{
  hipStream_t s = rt().stream;
  ProfScope ps("list_make_unique");
  const uint64_t words = (nvox3(L->dims) + 63) / 64;
  unsigned long long* bm = nullptr;
  uint32_t *pc = nullptr, *wpre = nullptr;
  GVX_CHECK(lu_buffers(words, &bm, &pc, &wpre));
  const unsigned gn = grid_for(n, 256, 4096), gw = grid_for(words, 256, 8192);
  hipLaunchKernelGGL(k_lu_mark, dim3(gn), dim3(256), 0, s, L->ids, keep, keep_from, n, bm);
  hipLaunchKernelGGL(k_lu_popc, dim3(gw), dim3(256), 0, s, bm, words, pc);
  GVX_HIP(hipGetLastError());
  GVX_CHECK(scan_u32(pc, wpre, words, wpre + words));
  GVX_CHECK(fill_u32(L->bits_t, 0u, n * 8));
  hipLaunchKernelGGL(k_lu_scatter, dim3(gn), dim3(256), 0, s, L->ids, L->coords, L->bits, keep, keep_from, n, bm, wpre,
                     L->ids_t, L->coords_t, L->bits_t);
  hipLaunchKernelGGL(k_lu_clear, dim3(gn), dim3(256), 0, s, L->ids, keep, keep_from, n, bm);
  // the unique count stays on the device for the copy back; the host reads it after
  // the call's one synchronisation (no wait in the middle of the call)
  hipLaunchKernelGGL(k_copy_entries, dim3(gn), dim3(256), 0, s, wpre + words, n, L->ids_t, L->coords_t, L->bits_t,
                     L->ids, L->coords, L->bits);
  GVX_HIP(hipGetLastError());
  GVX_HIP(hipMemcpyAsync(rt().host_small, wpre + words, 4, hipMemcpyDeviceToHost, s));
  GVX_CHECK(sync());
  L->n = *(const uint32_t*)rt().host_small;
  return GVX_OK;
}

// sort by id, OR duplicates, drop them
static int make_unique(ListS* L)
{
  const uint64_t n = L->n;
  if (n == 0) return GVX_OK;
  if (lu_suits(L, n)) {
    GVX_CHECK(tmp_reserve(L, n));
    return lu_unique(L, n, nullptr, n);
  }
  GVX_CHECK(tmp_reserve(L, n));
  hipStream_t s = rt().stream;
  ProfScope ps("list_make_unique");
  // keys -> ids_t (copy), perm -> v1
  GVX_HIP(hipMemcpyAsync(L->ids_t, L->ids, n * 4, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_iota, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, L->v1, n);
  GVX_HIP(hipGetLastError());
  GVX_CHECK(radix_sort(L, L->ids_t, L->v1, n));
  hipLaunchKernelGGL(k_heads, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, L->ids_t, n, L->flags);
  GVX_HIP(hipGetLastError());
  GVX_CHECK(scan_u32(L->flags, L->pos, n, L->pos + n));
  hipLaunchKernelGGL(k_unique, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, L->ids_t, L->v1, L->flags, L->pos, n,
                     L->coords, L->bits, L->ids, L->coords_t, L->bits_t);
  GVX_HIP(hipGetLastError());
  uint32_t m = 0;
  GVX_CHECK(read_u32(L->pos + n, &m));
  GVX_HIP(hipMemcpyAsync(L->coords, L->coords_t, (uint64_t)m * 12, hipMemcpyDeviceToDevice, s));
  GVX_HIP(hipMemcpyAsync(L->bits, L->bits_t, (uint64_t)m * 32, hipMemcpyDeviceToDevice, s));
  L->n = m;
  return sync();
}

// keep entries [base, base+cnt) whose flag is set (flags already in L->flags)
static int compact_tail(ListS* L, uint64_t base, uint64_t cnt)
{
  hipStream_t s = rt().stream;
  GVX_CHECK(scan_u32(L->flags, L->pos, cnt, L->pos + cnt));
  hipLaunchKernelGGL(k_compact, dim3(grid_for(cnt, 256, 4096)), dim3(256), 0, s, L->ids + base, L->coords + 3 * base,
                     L->bits + 8 * base, L->flags, L->pos, cnt, L->ids_t, L->coords_t, L->bits_t);
  GVX_HIP(hipGetLastError());
  uint32_t kept = 0;
  GVX_CHECK(read_u32(L->pos + cnt, &kept));
  GVX_HIP(hipMemcpyAsync(L->ids + base, L->ids_t, (uint64_t)kept * 4, hipMemcpyDeviceToDevice, s));
  GVX_HIP(hipMemcpyAsync(L->coords + 3 * base, L->coords_t, (uint64_t)kept * 12, hipMemcpyDeviceToDevice, s));
  GVX_HIP(hipMemcpyAsync(L->bits + 8 * base, L->bits_t, (uint64_t)kept * 32, hipMemcpyDeviceToDevice, s));
  L->n = base + kept;
  return GVX_OK;
}

// ---------------------------------------------------------------- merge kernels
// applyOffsetOperator + optional meaning fill (TemplateVoxelList.hpp:70-90, 878-916)
__global__ void k_append_list(const uint32_t* __restrict__ ids, const uint32_t* __restrict__ coords,
                              const uint32_t* __restrict__ bits, uint64_t n, int64_t addr_off, int32_t ox, int32_t oy,
                              int32_t oz, int meaning, uint32_t* __restrict__ oids, uint32_t* __restrict__ ocoords,
                              uint32_t* __restrict__ obits)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    oids[i] = (uint32_t)((int64_t)ids[i] + addr_off);
    ocoords[3 * i] = coords[3 * i] + (uint32_t)ox;
    ocoords[3 * i + 1] = coords[3 * i + 1] + (uint32_t)oy;
    ocoords[3 * i + 2] = coords[3 * i + 2] + (uint32_t)oz;
    uint4* ob = reinterpret_cast<uint4*>(obits + 8 * i);
    if (meaning >= 0) {
      uint32_t w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      w[meaning >> 5] = 1u << (meaning & 31);
      ob[0] = make_uint4(w[0], w[1], w[2], w[3]);
      ob[1] = make_uint4(w[4], w[5], w[6], w[7]);
    } else {
      const uint4* b = reinterpret_cast<const uint4*>(bits + 8 * i);
      ob[0] = b[0];
      ob[1] = b[1];
    }
  }
}

// occupied flags of a BitVector map (is_occupied, TemplateVoxelList.hpp:969-990)
__global__ void k_map_occ_flags(const uint32_t* __restrict__ m, uint64_t n, uint32_t* __restrict__ flags)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint4* p = reinterpret_cast<const uint4*>(m + 8 * i);
    flags[i] = bits_occupied(p[0], p[1]) ? 1u : 0u;
  }
}

__global__ void k_map_to_list(const uint32_t* __restrict__ m, uint64_t n, uint32_t dx, uint32_t dy,
                              const uint32_t* __restrict__ flags, const uint32_t* __restrict__ pos, int64_t addr_off,
                              int32_t ox, int32_t oy, int32_t oz, int meaning, uint32_t* __restrict__ oids,
                              uint32_t* __restrict__ ocoords, uint32_t* __restrict__ obits)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if (!flags[i]) continue;
    const uint64_t o = pos[i];
    const uint32_t x = (uint32_t)(i % dx), y = (uint32_t)((i / dx) % dy), z = (uint32_t)(i / ((uint64_t)dx * dy));
    oids[o] = (uint32_t)((int64_t)i + addr_off);
    ocoords[3 * o] = x + (uint32_t)ox;
    ocoords[3 * o + 1] = y + (uint32_t)oy;
    ocoords[3 * o + 2] = z + (uint32_t)oz;
    uint4* ob = reinterpret_cast<uint4*>(obits + 8 * o);
    if (meaning >= 0) {
      uint32_t w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      w[meaning >> 5] = 1u << (meaning & 31);
      ob[0] = make_uint4(w[0], w[1], w[2], w[3]);
      ob[1] = make_uint4(w[4], w[5], w[6], w[7]);
    } else {
      const uint4* p = reinterpret_cast<const uint4*>(m + 8 * i);
      ob[0] = p[0];
      ob[1] = p[1];
    }
  }
}

// ---------------------------------------------------------------- list x map / list x list
// VoxelListOperations.hpp:146-334.  mode 0 DefaultCollider, 1 isOccupied + OR of entry
// bits, 2 isOccupied + typemask.  Map voxel at id + addr_off, valid in [0, N).
template <int MT>
__global__ void __launch_bounds__(256) k_list_collide_map(const uint32_t* __restrict__ ids, uint32_t* __restrict__ bits,
                                                          uint64_t len, const uint8_t* __restrict__ map, uint64_t nvox,
                                                          int64_t addr_off, int thr_i, float thr_f, int mode,
                                                          uint4 tm0, uint4 tm1, unsigned long long* __restrict__ count,
                                                          uint32_t* __restrict__ mask)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t c = 0;
  uint32_t acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) {
    const int64_t j = (int64_t)ids[i] + addr_off;
    bool hit = false;
    if (j >= 0 && j < (int64_t)nvox) {
      bool map_occ;
      if constexpr (MT == GVX_MT_BITVECTOR_VOXELMAP) {
        const uint4* p = reinterpret_cast<const uint4*>(map + (uint64_t)j * 32);
        map_occ = bits_occupied(p[0], p[1]);
      } else {
        const int8_t o = (int8_t)map[j];
        map_occ = (mode == 0) ? ((int)o >= thr_i) : ((float)o >= thr_f);
      }
      if (map_occ) {
        uint4* e = reinterpret_cast<uint4*>(bits + 8 * i);
        uint4 e0 = e[0], e1 = e[1];
        if (mode == 0) hit = bits_occupied(e0, e1);
        else if (mode == 1) hit = true;
        else
          hit = ((e0.x & tm0.x) | (e0.y & tm0.y) | (e0.z & tm0.z) | (e0.w & tm0.w) | (e1.x & tm1.x) | (e1.y & tm1.y) |
                 (e1.z & tm1.z) | (e1.w & tm1.w)) != 0u;
        if (hit) {
          if (mode == 1) {
            acc[0] |= e0.x; acc[1] |= e0.y; acc[2] |= e0.z; acc[3] |= e0.w;
            acc[4] |= e1.x; acc[5] |= e1.y; acc[6] |= e1.z; acc[7] |= e1.w;
          }
          reinterpret_cast<uint32_t*>(e)[0] = e0.x | (1u << M_COLLISION);
        }
      }
    }
    c += __popcll(wave_ballot(hit));
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) acc[k] = wave_or(acc[k]);
  if ((threadIdx.x & 63) == 0) {
    if (c) atomicAdd(count, (unsigned long long)c);
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (acc[k]) atomicOr(mask + k, acc[k]);
  }
}

// binary search of A's ids in B (collideVoxellists / findMatchingVoxels)
__global__ void __launch_bounds__(256) k_list_collide_list(const uint32_t* __restrict__ ia,
                                                           const uint32_t* __restrict__ ba, uint64_t na,
                                                           const uint32_t* __restrict__ ib,
                                                           const uint32_t* __restrict__ bb, uint64_t nb,
                                                           unsigned long long* __restrict__ count,
                                                           uint32_t* __restrict__ mask)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t c = 0;
  uint32_t acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < na; i += stride) {
    const uint32_t key = ia[i];
    uint64_t lo = 0, hi = nb;
    while (lo < hi) {
      uint64_t mid = (lo + hi) >> 1;
      if (ib[mid] < key) lo = mid + 1; else hi = mid;
    }
    const bool hit = lo < nb && ib[lo] == key;
    if (hit) {
#pragma unroll
      for (int k = 0; k < 8; ++k) acc[k] |= ba[8 * i + k] | bb[8 * lo + k];
    }
    c += __popcll(wave_ballot(hit));
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) acc[k] = wave_or(acc[k]);
  if ((threadIdx.x & 63) == 0) {
    if (c) atomicAdd(count, (unsigned long long)c);
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (acc[k]) atomicOr(mask + k, acc[k]);
  }
}

// collideWithBitcheck (BitVoxelList.hpp:260-292): ids of A also in B whose
// voxels collide -- margin 0: any common bit of the full vectors
// (BitvectorCollision, BitVoxelList.h:40-48), else bitMarginCollisionCheck with
// sv_offset 0 (BitvectorCollisionWithBitshift, BitVoxelList.h:56-73)
__global__ void __launch_bounds__(256) k_list_bitcheck(const uint32_t* __restrict__ ia,
                                                       const uint32_t* __restrict__ ba, uint64_t na,
                                                       const uint32_t* __restrict__ ib,
                                                       const uint32_t* __restrict__ bb, uint64_t nb,
                                                       uint32_t margin, unsigned long long* __restrict__ count)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t c = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < na; i += stride) {
    const uint32_t key = ia[i];
    uint64_t lo = 0, hi = nb;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (ib[mid] < key) lo = mid + 1; else hi = mid;
    }
    bool hit = false;
    if (lo < nb && ib[lo] == key) {
      uint32_t va[8], vb[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) { va[k] = ba[8 * i + k]; vb[k] = bb[8 * lo + k]; }
      if (margin == 0) {
        uint32_t any = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) any |= va[k] & vb[k];
        hit = any != 0u;
      } else {
        uint32_t r[8];
        hit = bitmargin(va, vb, margin, 0, r);
      }
    }
    c += __popcll(wave_ballot(hit));
  }
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(count, (unsigned long long)c);
}

// collideCountingPerMeaning (BitVoxelList.hpp:295-343): every set bit of A's
// voxel whose id B also holds counts for its meaning; block-local counters in LDS
__global__ void __launch_bounds__(256) k_list_per_meaning(const uint32_t* __restrict__ ia,
                                                          const uint32_t* __restrict__ ba, uint64_t na,
                                                          const uint32_t* __restrict__ ib, uint64_t nb,
                                                          unsigned long long* __restrict__ counts)
{
  __shared__ uint32_t cnt[256];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < na; i += stride) {
    const uint32_t key = ia[i];
    uint64_t lo = 0, hi = nb;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (ib[mid] < key) lo = mid + 1; else hi = mid;
    }
    if (lo < nb && ib[lo] == key) {
      for (int k = 0; k < 8; ++k) {
        uint32_t w = ba[8 * i + k];
        while (w) {
          const uint32_t b = (uint32_t)__builtin_ctz(w);
          w &= w - 1u;
          atomicAdd(&cnt[32 * k + b], 1u);
        }
      }
    }
  }
  __syncthreads();
  if (cnt[threadIdx.x]) atomicAdd(counts + threadIdx.x, (unsigned long long)cnt[threadIdx.x]);
}

static bool dims_equal(const ListS* a, const ListS* b)
{
  return memcmp(a->dims, b->dims, sizeof(a->dims)) == 0;
}

static int64_t signed_index(const uint32_t d[3], const int32_t* o)
{
  if (!o) return 0;
  return (int64_t)o[2] * (int64_t)d[0] * (int64_t)d[1] + (int64_t)o[1] * (int64_t)d[0] + (int64_t)o[0];
}

// append points already on the device (cloud table on the device too), then
// remove_out_of_bounds and make_unique; does not touch the staging buffer
int list_insert_device(void* lv, const float* dp, uint64_t n, const uint32_t* ends, uint32_t ncl,
                       const uint32_t* meanings, uint32_t meaning)
{
  ListS* l = static_cast<ListS*>(static_cast<gvx_list_s*>(lv));
  if (n == 0) return GVX_OK;
  std::lock_guard<std::recursive_mutex> g(l->mu);
  std::lock_guard<std::recursive_mutex> g2(rt().mu);
  const uint64_t base = l->n;
  GVX_CHECK(list_reserve(l, base + n));
  GVX_CHECK(tmp_reserve(l, base + n));
  {
    ProfScope ps("list_insert_points");
    hipLaunchKernelGGL(k_list_points, dim3(grid_for(n, 256, 1u << 16)), dim3(256), 0, rt().stream, dp, n, l->dims[0],
                       l->dims[1], l->dims[2], l->side, ends, ncl, meanings, meaning, l->ids + base,
                       l->coords + 3 * base, l->bits + 8 * base, l->flags);
    GVX_HIP(hipGetLastError());
  }
  if (lu_suits(l, base + n)) return lu_unique(l, base + n, l->flags, base);  // out-of-map points: keep flag 0
  GVX_CHECK(compact_tail(l, base, n));
  return make_unique(l);
}

}  // namespace gvx

using namespace gvx;

static int check_list(gvx_list l)
{
  if (!l) return fail(GVX_ERR_BAD_ARG, "null list handle");
  return GVX_OK;
}

namespace gvx {
int shift_sv_async(uint8_t* voxels, uint64_t n, uint32_t shift);

// replace the contents with n host-side entries, taken as they are (a list dump
// holds a sorted, duplicate-free list)
int list_assign_host(gvx_list l, uint64_t n, const uint32_t* ids, const uint32_t* coords, const uint8_t* bits)
{
  GVX_CHECK(check_list(l));
  std::lock_guard<std::recursive_mutex> g(l->mu);
  GVX_CHECK(list_reserve(l, n));
  if (n) {
    GVX_CHECK(copy_in(l->ids, ids, n * 4, GVX_MEM_HOST));
    GVX_CHECK(copy_in(l->coords, coords, n * 12, GVX_MEM_HOST));
    GVX_CHECK(copy_in(l->bits, bits, n * 32, GVX_MEM_HOST));
  }
  l->n = n;
  l->ids_in_range = false;  // taken as they are
  return sync();
}

}  // namespace gvx

extern "C" {

int gvx_list_create(uint32_t dx, uint32_t dy, uint32_t dz, float side, gvx_list* out)
{
  if (!out) return fail(GVX_ERR_BAD_ARG, "null out");
  *out = nullptr;
  GVX_CHECK(ensure_init());
  if (dx == 0 || dy == 0 || dz == 0 || !(side > 0.0f)) return fail(GVX_ERR_BAD_ARG, "gvx_list_create: bad dims/side");
  gvx_list l = new gvx_list_s();
  l->dims[0] = dx; l->dims[1] = dy; l->dims[2] = dz;
  l->side = side;
  *out = l;
  return GVX_OK;
}

int gvx_list_destroy(gvx_list l)
{
  if (!l) return GVX_OK;
  (void)sync();
  uint32_t* bufs[] = {l->ids, l->coords, l->bits, l->k2, l->v1, l->v2, l->flags, l->pos, l->hist, l->ids_t,
                      l->coords_t, l->bits_t};
  for (auto b : bufs) dfree(b);
  delete l;
  return GVX_OK;
}

int gvx_list_size(gvx_list l, uint64_t* n)
{
  GVX_CHECK(check_list(l));
  if (!n) return fail(GVX_ERR_BAD_ARG, "null");
  *n = l->n;
  return GVX_OK;
}

int gvx_list_info(gvx_list l, uint32_t dims[3], float* side)
{
  GVX_CHECK(check_list(l));
  if (dims) { dims[0] = l->dims[0]; dims[1] = l->dims[1]; dims[2] = l->dims[2]; }
  if (side) *side = l->side;
  return GVX_OK;
}

int gvx_list_clear(gvx_list l)
{
  GVX_CHECK(check_list(l));
  std::lock_guard<std::recursive_mutex> g(l->mu);
  l->n = 0;
  l->ids_in_range = true;
  return GVX_OK;
}

int gvx_list_read(gvx_list l, uint32_t* ids, uint32_t* coords, uint8_t* bits, int mem)
{
  GVX_CHECK(check_list(l));
  std::lock_guard<std::recursive_mutex> g(l->mu);
  if (ids) GVX_CHECK(copy_out(ids, l->ids, l->n * 4, mem));
  if (coords) GVX_CHECK(copy_out(coords, l->coords, l->n * 12, mem));
  if (bits) GVX_CHECK(copy_out(bits, l->bits, l->n * 32, mem));
  return sync();
}

int gvx_list_device_ptrs(gvx_list l, uint32_t** ids, uint32_t** coords, uint8_t** bits)
{
  GVX_CHECK(check_list(l));
  if (ids) *ids = l->ids;
  if (coords) *coords = l->coords;
  if (bits) *bits = (uint8_t*)l->bits;
  return GVX_OK;
}

int gvx_list_insert_points(gvx_list l, const float* xyz, uint64_t n, const uint32_t* sizes, uint32_t ncl,
                           const uint32_t* meanings, uint32_t meaning, int mem)
{
  GVX_CHECK(check_list(l));
  if (n && !xyz) return fail(GVX_ERR_BAD_ARG, "null points");
  if (n > 0xFFFFFFFFull) return fail(GVX_ERR_UNSUPPORTED, "more than 2^32 points");
  std::vector<uint32_t> tab;
  if (sizes) {
    if (ncl == 0) return fail(GVX_ERR_BAD_ARG, "cloud sizes without clouds");
    if (!meanings) return fail(GVX_ERR_BAD_ARG, "per-cloud meanings required with cloud sizes");
    tab.resize(2 * (size_t)ncl);
    uint64_t tot = 0;
    for (uint32_t c = 0; c < ncl; ++c) {
      tot += sizes[c];
      tab[c] = (uint32_t)tot;
      if (meanings[c] > 255) return fail(GVX_ERR_BAD_ARG, "meaning > 255");
      tab[ncl + c] = meanings[c];
    }
    if (tot != n) return fail(GVX_ERR_BAD_ARG, "cloud sizes do not sum to n");
  } else if (meaning > 255) {
    return fail(GVX_ERR_BAD_ARG, "meaning > 255");
  }
  if (n == 0) return GVX_OK;
  std::lock_guard<std::recursive_mutex> g2(rt().mu);
  size_t pb = (n * 12 + 255) & ~(size_t)255;
  void* st = nullptr;
  GVX_CHECK(stage_reserve(pb + tab.size() * 4 + 4, &st));
  const float* dp = xyz;
  if (mem == GVX_MEM_HOST) {
    GVX_HIP(hipMemcpyAsync(st, xyz, n * 12, hipMemcpyHostToDevice, rt().stream));
    dp = (const float*)st;
  }
  uint32_t* dtab = nullptr;
  if (!tab.empty()) {
    dtab = (uint32_t*)((char*)st + pb);
    GVX_HIP(hipMemcpyAsync(dtab, tab.data(), tab.size() * 4, hipMemcpyHostToDevice, rt().stream));
  }
  return gvx::list_insert_device(l, dp, n, dtab, ncl, dtab ? dtab + ncl : nullptr, meaning);
}

int gvx_list_insert_coords(gvx_list l, const uint32_t* xyz, uint64_t n, uint32_t meaning, int mem)
{
  GVX_CHECK(check_list(l));
  if (n && !xyz) return fail(GVX_ERR_BAD_ARG, "null coords");
  if (meaning > 255) return fail(GVX_ERR_BAD_ARG, "meaning > 255");
  if (n == 0) return GVX_OK;
  std::lock_guard<std::recursive_mutex> g(l->mu);
  std::lock_guard<std::recursive_mutex> g2(rt().mu);
  const uint64_t base = l->n;
  GVX_CHECK(list_reserve(l, base + n));
  GVX_CHECK(tmp_reserve(l, base + n));
  const void* d = nullptr;
  GVX_CHECK(stage_input(xyz, n * 12, mem, &d));
  hipLaunchKernelGGL(k_list_coords, dim3(grid_for(n, 256, 1u << 16)), dim3(256), 0, rt().stream, (const uint32_t*)d,
                     n, l->dims[0], l->dims[1], l->dims[2], meaning, l->ids + base, l->coords + 3 * base,
                     l->bits + 8 * base, l->flags);
  GVX_HIP(hipGetLastError());
  if (lu_suits(l, base + n)) return lu_unique(l, base + n, l->flags, base);
  GVX_CHECK(compact_tail(l, base, n));
  return make_unique(l);
}

int gvx_list_merge_list(gvx_list self, gvx_list other, const int32_t offset[3], int meaning)
{
  GVX_CHECK(check_list(self));
  GVX_CHECK(check_list(other));
  if (meaning > 255) return fail(GVX_ERR_BAD_ARG, "meaning > 255");
  if (self == other) return fail(GVX_ERR_BAD_ARG, "merging a list into itself");
  std::unique_lock<std::recursive_mutex> g1(self->mu, std::defer_lock), g2(other->mu, std::defer_lock);
  std::lock(g1, g2);
  std::lock_guard<std::recursive_mutex> g3(rt().mu);
  const uint64_t base = self->n, on = other->n;
  if (on == 0) return GVX_OK;
  GVX_CHECK(list_reserve(self, base + on));
  GVX_CHECK(tmp_reserve(self, base + on));
  hipLaunchKernelGGL(k_append_list, dim3(grid_for(on, 256, 4096)), dim3(256), 0, rt().stream, other->ids, other->coords,
                     other->bits, on, signed_index(self->dims, offset), offset ? offset[0] : 0,
                     offset ? offset[1] : 0, offset ? offset[2] : 0, meaning, self->ids + base,
                     self->coords + 3 * base, self->bits + 8 * base);
  GVX_HIP(hipGetLastError());
  self->n = base + on;
  const bool no_off = !offset || (offset[0] == 0 && offset[1] == 0 && offset[2] == 0);
  self->ids_in_range = self->ids_in_range && other->ids_in_range && no_off && dims_equal(self, other);
  return make_unique(self);
}

int gvx_list_merge_map(gvx_list self, gvx_map map, const int32_t offset[3], int meaning)
{
  GVX_CHECK(check_list(self));
  if (!map) return fail(GVX_ERR_BAD_ARG, "null map");
  if (map->type != GVX_MT_BITVECTOR_VOXELMAP)
    return fail(GVX_ERR_UNSUPPORTED, "THIS OPERATION IS NOT YET SUPPORTED BY THE DATA STRUCTURE!");
  if (meaning > 255) return fail(GVX_ERR_BAD_ARG, "meaning > 255");
  std::unique_lock<std::recursive_mutex> g1(self->mu, std::defer_lock), g2(map->mu, std::defer_lock);
  std::lock(g1, g2);
  std::lock_guard<std::recursive_mutex> g3(rt().mu);
  const uint64_t nv = map->n;
  uint32_t *flags = nullptr, *pos = nullptr;
  GVX_CHECK(dalloc((void**)&flags, nv * 4));
  int s = dalloc((void**)&pos, (nv + 1) * 4);
  if (s != GVX_OK) { dfree(flags); return s; }
  auto cleanup = [&]() { (void)sync(); dfree(flags); dfree(pos); };
  hipLaunchKernelGGL(k_map_occ_flags, dim3(grid_for(nv, 256, 8192)), dim3(256), 0, rt().stream,
                     (const uint32_t*)map->data, nv, flags);
  s = scan_u32(flags, pos, nv, pos + nv);
  uint32_t occ = 0;
  if (s == GVX_OK) s = read_u32(pos + nv, &occ);
  const uint64_t base = self->n;
  if (s == GVX_OK) s = list_reserve(self, base + occ);
  if (s == GVX_OK) s = tmp_reserve(self, base + occ);
  if (s == GVX_OK) {
    hipLaunchKernelGGL(k_map_to_list, dim3(grid_for(nv, 256, 8192)), dim3(256), 0, rt().stream,
                       (const uint32_t*)map->data, nv, map->dims[0], map->dims[1], flags, pos,
                       signed_index(self->dims, offset), offset ? offset[0] : 0, offset ? offset[1] : 0,
                       offset ? offset[2] : 0, meaning, self->ids + base, self->coords + 3 * base,
                       self->bits + 8 * base);
    if (hipGetLastError() != hipSuccess) s = fail(GVX_ERR_HIP, "k_map_to_list launch");
  }
  cleanup();
  if (s != GVX_OK) return s;
  self->n = base + occ;
  const bool no_off = !offset || (offset[0] == 0 && offset[1] == 0 && offset[2] == 0);
  self->ids_in_range = self->ids_in_range && no_off && memcmp(self->dims, map->dims, sizeof(self->dims)) == 0;
  return make_unique(self);
}

int gvx_list_collide_map(gvx_list self, gvx_map map, float threshold, const int32_t offset[3], int mode,
                         const uint8_t typemask[32], uint64_t* count, uint8_t mask_out[32])
{
  GVX_CHECK(check_list(self));
  if (!map || !count) return fail(GVX_ERR_BAD_ARG, "null map/count");
  if (mode < 0 || mode > 2) return fail(GVX_ERR_BAD_ARG, "bad mode");
  if (mode == 2 && !typemask) return fail(GVX_ERR_BAD_ARG, "typemask required in mode 2");
  if (memcmp(self->dims, map->dims, sizeof(self->dims)) != 0)
    return fail(GVX_ERR_DIM_MISMATCH,
                "The dimensions of the Voxellist reference map do not match the colliding voxel map dimensions.");
  if (map->type == GVX_MT_DISTANCE_VOXELMAP) {
    *count = 0;
    if (mask_out) memset(mask_out, 0, 32);
    return GVX_OK;
  }
  std::unique_lock<std::recursive_mutex> g1(self->mu, std::defer_lock), g2(map->mu, std::defer_lock);
  std::lock(g1, g2);
  std::lock_guard<std::recursive_mutex> g3(rt().mu);
  uint32_t tm[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (typemask) memcpy(tm, typemask, 32);
  unsigned long long* dcount = (unsigned long long*)rt().dev_small;
  uint32_t* dmask = (uint32_t*)rt().dev_small + 2;
  GVX_CHECK(fill_u32(rt().dev_small, 0u, 12));
  if (self->n) {
    ProfScope ps("list_collide_map");
    unsigned g = grid_for(self->n, 256, 8192);
    const int64_t ao = signed_index(self->dims, offset);
    const int ti = float_to_probability(threshold);
    if (map->type == GVX_MT_BITVECTOR_VOXELMAP)
      hipLaunchKernelGGL(k_list_collide_map<GVX_MT_BITVECTOR_VOXELMAP>, dim3(g), dim3(256), 0, rt().stream, self->ids,
                         self->bits, self->n, (const uint8_t*)map->data, map->n, ao, ti, threshold, mode,
                         make_uint4(tm[0], tm[1], tm[2], tm[3]), make_uint4(tm[4], tm[5], tm[6], tm[7]), dcount,
                         dmask);
    else
      hipLaunchKernelGGL(k_list_collide_map<GVX_MT_PROBAB_VOXELMAP>, dim3(g), dim3(256), 0, rt().stream, self->ids,
                         self->bits, self->n, (const uint8_t*)map->data, map->n, ao, ti, threshold, mode,
                         make_uint4(tm[0], tm[1], tm[2], tm[3]), make_uint4(tm[4], tm[5], tm[6], tm[7]), dcount,
                         dmask);
    GVX_HIP(hipGetLastError());
  }
  GVX_HIP(hipMemcpyAsync(rt().host_small, rt().dev_small, 40, hipMemcpyDeviceToHost, rt().stream));
  GVX_CHECK(sync());
  *count = *(unsigned long long*)rt().host_small;
  if (mask_out) memcpy(mask_out, (uint8_t*)rt().host_small + 8, 32);
  return GVX_OK;
}

int gvx_list_collide_list(gvx_list self, gvx_list other, uint64_t* count, uint8_t mask_out[32])
{
  GVX_CHECK(check_list(self));
  GVX_CHECK(check_list(other));
  if (!count) return fail(GVX_ERR_BAD_ARG, "null count");
  std::unique_lock<std::recursive_mutex> g1(self->mu, std::defer_lock), g2(other->mu, std::defer_lock);
  if (self == other) g1.lock(); else std::lock(g1, g2);
  std::lock_guard<std::recursive_mutex> g3(rt().mu);
  unsigned long long* dcount = (unsigned long long*)rt().dev_small;
  uint32_t* dmask = (uint32_t*)rt().dev_small + 2;
  GVX_CHECK(fill_u32(rt().dev_small, 0u, 12));
  if (self->n && other->n) {
    hipLaunchKernelGGL(k_list_collide_list, dim3(grid_for(self->n, 256, 8192)), dim3(256), 0, rt().stream, self->ids,
                       self->bits, self->n, other->ids, other->bits, other->n, dcount, dmask);
    GVX_HIP(hipGetLastError());
  }
  GVX_HIP(hipMemcpyAsync(rt().host_small, rt().dev_small, 40, hipMemcpyDeviceToHost, rt().stream));
  GVX_CHECK(sync());
  *count = *(unsigned long long*)rt().host_small;
  if (mask_out) memcpy(mask_out, (uint8_t*)rt().host_small + 8, 32);
  return GVX_OK;
}

int gvx_list_shift_sv(gvx_list l, uint32_t shift)
{
  GVX_CHECK(check_list(l));
  if (shift > 56) return fail(GVX_ERR_BAD_ARG, "Maximum shift size is 56");
  std::lock_guard<std::recursive_mutex> g(l->mu);
  GVX_CHECK(shift_sv_async((uint8_t*)l->bits, l->n, shift));
  return sync();
}

int gvx_list_collide_bitcheck(gvx_list self, gvx_list other, uint32_t margin, uint64_t* count)
{
  GVX_CHECK(check_list(self));
  GVX_CHECK(check_list(other));
  if (!count) return fail(GVX_ERR_BAD_ARG, "null count");
  if (margin > 31) return fail(GVX_ERR_BAD_ARG, "bit margin must be <= 31");
  std::unique_lock<std::recursive_mutex> g1(self->mu, std::defer_lock), g2(other->mu, std::defer_lock);
  if (self == other) g1.lock(); else std::lock(g1, g2);
  std::lock_guard<std::recursive_mutex> g3(rt().mu);
  unsigned long long* dcount = (unsigned long long*)rt().dev_small;
  GVX_CHECK(fill_u32(rt().dev_small, 0u, 2));
  if (self->n && other->n) {
    hipLaunchKernelGGL(k_list_bitcheck, dim3(grid_for(self->n, 256, 8192)), dim3(256), 0, rt().stream, self->ids,
                       self->bits, self->n, other->ids, other->bits, other->n, margin, dcount);
    GVX_HIP(hipGetLastError());
  }
  GVX_HIP(hipMemcpyAsync(rt().host_small, rt().dev_small, 8, hipMemcpyDeviceToHost, rt().stream));
  GVX_CHECK(sync());
  *count = *(unsigned long long*)rt().host_small;
  return GVX_OK;
}

int gvx_list_collide_per_meaning(gvx_list self, gvx_list other, uint64_t counts[256], uint64_t* total)
{
  GVX_CHECK(check_list(self));
  GVX_CHECK(check_list(other));
  if (!counts) return fail(GVX_ERR_BAD_ARG, "null counts");
  std::unique_lock<std::recursive_mutex> g1(self->mu, std::defer_lock), g2(other->mu, std::defer_lock);
  if (self == other) g1.lock(); else std::lock(g1, g2);
  std::lock_guard<std::recursive_mutex> g3(rt().mu);
  void* dc = nullptr;
  GVX_CHECK(stage_reserve(256 * 8, &dc));
  GVX_CHECK(fill_u32(dc, 0u, 512));
  if (self->n && other->n) {
    hipLaunchKernelGGL(k_list_per_meaning, dim3(grid_for(self->n, 256, 4096)), dim3(256), 0, rt().stream, self->ids,
                       self->bits, self->n, other->ids, other->n, (unsigned long long*)dc);
    GVX_HIP(hipGetLastError());
  }
  GVX_CHECK(copy_out(counts, dc, 256 * 8, GVX_MEM_HOST));
  GVX_CHECK(sync());
  uint64_t t = 0;
  for (int k = 0; k < 256; ++k) t += counts[k];
  if (total) *total = t;
  return GVX_OK;
}

}  // extern "C"
