// gvx_maps.hip -- dense voxel maps: create/clear, voxelization, map x map collision.
#include "gvx_internal.hpp"

namespace gvx {

constexpr int TBIT = GVX_MT_BITVECTOR_VOXELMAP, TPROB = GVX_MT_PROBAB_VOXELMAP, TDIST = GVX_MT_DISTANCE_VOXELMAP;

// ---------------------------------------------------------------- voxel updates
template <int T>
__device__ __forceinline__ void voxel_insert(uint8_t* map, uint64_t idx, uint32_t x, uint32_t y, uint32_t z,
                                             uint32_t meaning, uint32_t uninit)
{
  if constexpr (T == TBIT) {
    // BitVoxel::insert = setBit (BitVoxel.hpp:60-65); atomic OR into the word holding the bit
    atomicOr(reinterpret_cast<uint32_t*>(map) + idx * 8 + (meaning >> 5), 1u << (meaning & 31));
  } else if constexpr (T == TPROB) {
    // ProbabilisticVoxel::insert (ProbabilisticVoxel.hpp:79-101) as a CAS on the 32-bit word
    uint32_t* w = reinterpret_cast<uint32_t*>(map + (idx & ~(uint64_t)3));
    const uint32_t sh = (uint32_t)(idx & 3) * 8;
    uint32_t old = __atomic_load_n(w, __ATOMIC_RELAXED);
    for (;;) {
      int cur = (int)(int8_t)((old >> sh) & 0xFF);
      int nv = prob_insert(cur, meaning);
      if (nv == cur) break;
      uint32_t neu = (old & ~(0xFFu << sh)) | ((uint32_t)(uint8_t)(int8_t)nv << sh);
      uint32_t prev = atomicCAS(w, old, neu);
      if (prev == old) break;
      old = prev;
    }
  } else {
    // DistanceVoxel::insert (DistanceVoxel.hpp:128-139)
    reinterpret_cast<uint32_t*>(map)[idx] = (meaning == M_OCCUPIED) ? (x | (y << 10) | (z << 20)) : uninit;
  }
}

// `count` ProbabilisticVoxel::insert(meaning) updates in a row on voxel idx, one CAS
// (identical updates: any order gives the same byte)
__device__ __forceinline__ void prob_insert_n(uint8_t* map, uint64_t idx, uint32_t meaning, uint32_t count)
{
  uint32_t* w = reinterpret_cast<uint32_t*>(map + (idx & ~(uint64_t)3));
  const uint32_t sh = (uint32_t)(idx & 3) * 8;
  uint32_t old = __atomic_load_n(w, __ATOMIC_RELAXED);
  for (;;) {
    const int cur = (int)(int8_t)((old >> sh) & 0xFF);
    int nv = cur;
    for (uint32_t k = 0; k < count; ++k) {
      const int n2 = prob_insert(nv, meaning);
      if (n2 == nv) break;
      nv = n2;
    }
    if (nv == cur) return;
    const uint32_t neu = (old & ~(0xFFu << sh)) | ((uint32_t)(uint8_t)(int8_t)nv << sh);
    const uint32_t prev = atomicCAS(w, old, neu);
    if (prev == old) return;
    old = prev;
  }
}

// points: kernelInsertGlobalPointCloud / kernelInsertMetaPointCloud (VoxelMapOperations.hpp:241-575).
// cloud_ends: inclusive prefix of the cloud sizes (n_clouds entries) or null.
template <int T>
__global__ void __launch_bounds__(256) k_insert_points(uint8_t* __restrict__ map, uint32_t dx, uint32_t dy, uint32_t dz,
                                                       float side, const float* __restrict__ pts, uint64_t n,
                                                       const uint32_t* __restrict__ cloud_ends, uint32_t n_clouds,
                                                       const uint32_t* __restrict__ meanings, uint32_t meaning,
                                                       uint32_t uninit, int* __restrict__ outside)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    uint32_t m = meaning;
    if (cloud_ends) {
      uint32_t lo = 0, hi = n_clouds - 1;
      while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if ((uint64_t)cloud_ends[mid] > i) hi = mid; else lo = mid + 1;
      }
      m = meanings[lo];
    }
    const uint32_t x = quantize(pts[3 * i], side), y = quantize(pts[3 * i + 1], side),
                   z = quantize(pts[3 * i + 2], side);
    const bool in = (x < dx) && (y < dy) && (z < dz);
    if constexpr (T == TPROB) {
      // neighbouring points often share a voxel: a run of lanes on one voxel
      // updates it with one CAS (identical updates commute)
      const uint64_t idx = in ? ((uint64_t)z * dy + y) * dx + x : 0ull;
      const uint32_t run = wave_run_head(in, idx, m);
      if (run) prob_insert_n(map, idx, m, run);
    } else if constexpr (T == TBIT) {
      // setting a bit is idempotent: one lane per run on one (voxel, meaning) sets it
      const uint64_t idx = in ? ((uint64_t)z * dy + y) * dx + x : 0ull;
      if (wave_run_head(in, idx, m)) voxel_insert<T>(map, idx, x, y, z, m, uninit);
    } else {
      if (in) voxel_insert<T>(map, ((uint64_t)z * dy + y) * dx + x, x, y, z, m, uninit);
    }
    if (wave_ballot(!in) && outside) *outside = 1;
  }
}

template <int T>
__global__ void __launch_bounds__(256) k_insert_coords(uint8_t* __restrict__ map, uint32_t dx, uint32_t dy, uint32_t dz,
                                                       const uint32_t* __restrict__ c, uint64_t n, uint32_t meaning,
                                                       uint32_t uninit, int* __restrict__ outside)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t x = c[3 * i], y = c[3 * i + 1], z = c[3 * i + 2];
    const bool in = (x < dx) && (y < dy) && (z < dz);
    if (in) voxel_insert<T>(map, ((uint64_t)z * dy + y) * dx + x, x, y, z, meaning, uninit);
    if (wave_ballot(!in) && outside) *outside = 1;
  }
}

// ---------------------------------------------------------------- collision
template <int T>
__device__ __forceinline__ bool occupied(const uint8_t* __restrict__ m, uint64_t i, int thr)
{
  if constexpr (T == TBIT) {
    const uint4* p = reinterpret_cast<const uint4*>(m + i * 32);
    return bits_occupied(p[0], p[1]);
  } else {
    return (int)(int8_t)m[i] >= thr;
  }
}

template <int T>
__device__ __forceinline__ void mark_collision(uint8_t* __restrict__ m, uint64_t i)
{
  if constexpr (T == TBIT) {
    reinterpret_cast<uint32_t*>(m)[i * 8] |= (1u << M_COLLISION);
  } else {
    m[i] = (uint8_t)(int8_t)prob_update((int)(int8_t)m[i], SENSOR_OCC);
  }
}

// kernelCollideVoxelMapsDebug (VoxelMapOperations.hpp:129-184) with DefaultCollider:
// a[i + off] against b[i] for i in [lo, hi); 64-bit count, wave ballot + popcount.
template <int TA, int TB>
__global__ void __launch_bounds__(256) k_collide(uint8_t* __restrict__ a, const uint8_t* __restrict__ b, uint64_t lo,
                                                 uint64_t hi, int64_t off, int thr,
                                                 unsigned long long* __restrict__ count)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t c = 0;
  for (uint64_t i = lo + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < hi; i += stride) {
    const uint64_t j = (uint64_t)((int64_t)i + off);
    const bool hit = occupied<TB>(b, i, thr) && occupied<TA>(a, j, thr);
    if (hit) mark_collision<TA>(a, j);
    c += __popcll(wave_ballot(hit));
  }
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(count, (unsigned long long)c);
}

constexpr int kLoadsInFlight = 4;  // per lane, in the kernels that stream a ProbVoxelMap

// Bit 7 of each byte set where that byte of x >= the byte of t (unsigned; t holds
// one value in all four bytes).  Where the high bits differ they decide; where they
// agree, (x | 0x80) - (t & 0x7F) carries the comparison of the low 7 bits into
// bit 7 without borrowing from the next byte.
__device__ __forceinline__ uint32_t bytes_ge(uint32_t x, uint32_t t)
{
  const uint32_t H = 0x80808080u;
  const uint32_t d = (x | H) - (t & ~H);
  return ((x & ~t) | (~(x ^ t) & d)) & H;
}

// Pass flags (bit 7 of each byte) of 16 int8 log-odds against thr: signed bytes
// compare as unsigned ones once their sign bits are flipped.
__device__ __forceinline__ void prob_pass(const uint4& v, int thr, uint32_t f[4])
{
  const uint32_t t = ((uint32_t)(uint8_t)thr ^ 0x80u) * 0x01010101u, s = 0x80808080u;
  f[0] = bytes_ge(v.x ^ s, t);
  f[1] = bytes_ge(v.y ^ s, t);
  f[2] = bytes_ge(v.z ^ s, t);
  f[3] = bytes_ge(v.w ^ s, t);
}

// The passing voxels of a wave's 64 lanes x 16 voxels, spread over its lanes: each
// lane lists its own (offset lane * 16 + voxel) in the wave's LDS list at its prefix
// position, then lane t visits entries t, t + 64, ...  A lane whose 16 voxels all
// pass (a surface along x) would otherwise read the other side 16 times in a row.
// Called by every lane of the wave (wave-uniform control); sq holds 1024 entries.
template <typename Visit>
__device__ __forceinline__ void wave_spread(const uint32_t f[4], uint16_t* sq, Visit&& visit)
{
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t cnt = __popc(f[0]) + __popc(f[1]) + __popc(f[2]) + __popc(f[3]);
  uint32_t incl = cnt;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t t = __shfl_up(incl, d, 64);
    if (lane >= (uint32_t)d) incl += t;
  }
  const uint32_t total = __shfl(incl, 63, 64);
  uint32_t pos = incl - cnt;
  for (int q = 0; q < 4; ++q)
    for (uint32_t m = f[q]; m != 0u; m &= m - 1u) sq[pos++] = (uint16_t)(lane * 16 + q * 4 + (__builtin_ctz(m) >> 3));
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (uint32_t t = lane; t < total; t += 64) visit((uint32_t)sq[t]);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();  // the list is rewritten by the next call
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The same collision when one side is a ProbVoxelMap: a lane takes 16 consecutive
// Prob voxels with one 16-byte load and reads the other side's voxel only where
// the Prob voxel passes the threshold (a collision needs both), so a sparse
// environment costs 1 B per voxel instead of 33.  KEY_A: the Prob side is self
// (a[j], j = i + off), else other (b[i]).  Chunks are aligned on the Prob side;
// map allocations are padded to 256 B, and voxels outside the valid range are
// masked off.
template <int TA, int TB, bool KEY_A>
__global__ void __launch_bounds__(256) k_collide_prob(uint8_t* __restrict__ a, const uint8_t* __restrict__ b,
                                                      uint64_t lo, uint64_t hi, int64_t off, int thr,
                                                      unsigned long long* __restrict__ count)
{
  const uint8_t* key = KEY_A ? a : b;
  const int64_t shift = KEY_A ? off : 0;  // key index = i + shift
  const uint64_t klo = (uint64_t)((int64_t)lo + shift), khi = (uint64_t)((int64_t)hi + shift);
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint32_t lane = threadIdx.x & 63;
  __shared__ uint16_t sq[4][1024];
  uint64_t c = 0;
  // kLoadsInFlight 16-byte loads are issued before any is used; the loop runs while
  // the wave's first lane is in range, so control stays wave-uniform
  for (uint64_t k0 = klo / 16 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; (k0 - lane) * 16 < khi;
       k0 += stride * kLoadsInFlight) {
    uint4 v[kLoadsInFlight];
#pragma unroll
    for (int u = 0; u < kLoadsInFlight; ++u) {
      const uint64_t k = k0 + (uint64_t)u * stride;
      v[u] = k * 16 < khi ? reinterpret_cast<const uint4*>(key)[k] : make_uint4(0u, 0u, 0u, 0u);
    }
#pragma unroll
    for (int u = 0; u < kLoadsInFlight; ++u) {
      const uint64_t k = k0 + (uint64_t)u * stride, kw = (k - lane) * 16;  // the wave's first voxel
      if (kw >= khi) break;
      uint32_t f[4] = {0u, 0u, 0u, 0u};
      if (k * 16 < khi) prob_pass(v[u], thr, f);
      if (wave_ballot((f[0] | f[1] | f[2] | f[3]) != 0u) == 0ull) continue;
      wave_spread(f, sq[threadIdx.x >> 6], [&](uint32_t o) {
        const uint64_t kj = kw + o;
        if (kj < klo || kj >= khi) return;
        const uint64_t i = (uint64_t)((int64_t)kj - shift), j = (uint64_t)((int64_t)i + off);
        const bool hit = KEY_A ? occupied<TB>(b, i, thr) : occupied<TA>(a, j, thr);
        if (hit) {
          mark_collision<TA>(a, j);
          ++c;
        }
      });
    }
  }
  c = wave_sum_u64(c);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(count, (unsigned long long)c);
}

// kernelCollideVoxelMapsBitvector (VoxelMapOperations.hpp:187-239) with SVCollider,
// Bit x Bit: bitMarginCollisionCheck with margin 0 = a & b restricted to bits 4..255,
// mask |= a & b.  Masks are ORed over all colliding voxels.
__global__ void __launch_bounds__(256) k_collide_types(uint8_t* __restrict__ a, const uint8_t* __restrict__ b, uint64_t n,
                                                       unsigned long long* __restrict__ count,
                                                       uint32_t* __restrict__ mask)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t c = 0;
  uint32_t acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    uint4* pa = reinterpret_cast<uint4*>(a + i * 32);
    const uint4 a0 = pa[0], a1 = pa[1];
    // the other voxel is read only when self has a bit the AND can keep
    if (((a0.x & 0xFFFFFFF0u) | a0.y | a0.z | a0.w | a1.x | a1.y | a1.z | a1.w) == 0u) continue;
    const uint4* pb = reinterpret_cast<const uint4*>(b + i * 32);
    const uint4 b0 = pb[0], b1 = pb[1];
    const uint32_t r[8] = {a0.x & b0.x & 0xFFFFFFF0u, a0.y & b0.y, a0.z & b0.z, a0.w & b0.w,
                           a1.x & b1.x, a1.y & b1.y, a1.z & b1.z, a1.w & b1.w};
    if ((r[0] | r[1] | r[2] | r[3] | r[4] | r[5] | r[6] | r[7]) != 0u) {
#pragma unroll
      for (int k = 0; k < 8; ++k) acc[k] |= r[k];
      reinterpret_cast<uint32_t*>(pa)[0] = a0.x | (1u << M_COLLISION);
      ++c;
    }
  }
  c = wave_sum_u64(c);
#pragma unroll
  for (int k = 0; k < 8; ++k) acc[k] = wave_or(acc[k]);
  if ((threadIdx.x & 63) == 0) {
    if (c) atomicAdd(count, (unsigned long long)c);
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (acc[k]) atomicOr(mask + k, acc[k]);
  }
}

// Bit x Prob of the above: prob >= thr and a occupied, mask |= a.  The Prob side
// is read 16 voxels per lane; the bit voxel only where the Prob voxel passes.
__global__ void __launch_bounds__(256) k_collide_types_prob(uint8_t* __restrict__ a, const uint8_t* __restrict__ b,
                                                            uint64_t n, int thr, unsigned long long* __restrict__ count,
                                                            uint32_t* __restrict__ mask)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t c = 0;
  uint32_t acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const uint32_t lane = threadIdx.x & 63;
  __shared__ uint16_t sq[4][1024];
  for (uint64_t k0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; (k0 - lane) * 16 < n;
       k0 += stride * kLoadsInFlight) {
    uint4 v[kLoadsInFlight];
#pragma unroll
    for (int u = 0; u < kLoadsInFlight; ++u) {
      const uint64_t k = k0 + (uint64_t)u * stride;
      v[u] = k * 16 < n ? reinterpret_cast<const uint4*>(b)[k] : make_uint4(0u, 0u, 0u, 0u);
    }
#pragma unroll
    for (int u = 0; u < kLoadsInFlight; ++u) {
      const uint64_t k = k0 + (uint64_t)u * stride, kw = (k - lane) * 16;  // the wave's first voxel
      if (kw >= n) break;
      uint32_t f[4] = {0u, 0u, 0u, 0u};
      if (k * 16 < n) prob_pass(v[u], thr, f);
      if (wave_ballot((f[0] | f[1] | f[2] | f[3]) != 0u) == 0ull) continue;
      wave_spread(f, sq[threadIdx.x >> 6], [&](uint32_t o) {
        const uint64_t i = kw + o;
        if (i >= n) return;
        uint4* pa = reinterpret_cast<uint4*>(a + i * 32);
        const uint4 a0 = pa[0], a1 = pa[1];
        if (bits_occupied(a0, a1)) {
          acc[0] |= a0.x; acc[1] |= a0.y; acc[2] |= a0.z; acc[3] |= a0.w;
          acc[4] |= a1.x; acc[5] |= a1.y; acc[6] |= a1.z; acc[7] |= a1.w;
          reinterpret_cast<uint32_t*>(pa)[0] = a0.x | (1u << M_COLLISION);
          ++c;
        }
      });
    }
  }
  c = wave_sum_u64(c);
#pragma unroll
  for (int k = 0; k < 8; ++k) acc[k] = wave_or(acc[k]);
  if ((threadIdx.x & 63) == 0) {
    if (c) atomicAdd(count, (unsigned long long)c);
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (acc[k]) atomicOr(mask + k, acc[k]);
  }
}

// insertMetaPointCloudWithSelfCollisionCheck, one sub-cloud
// (kernelInsertMetaPointCloudSelfCollCheck, VoxelMapOperations.hpp:582-638).  Test
// pass: every point reads its voxel as it was before the sub-cloud; masked = voxel
// & mask colliding (a bit other than eBVM_FREE) flags the point and is ORed, with
// the sub-cloud's meaning, into acc[0..7] (one atomic per word per wave); acc[8]
// is the points-outside flag.
__global__ void __launch_bounds__(256) k_selfcoll_test(const uint32_t* __restrict__ map, uint32_t dx, uint32_t dy,
                                                       uint32_t dz, float side, const float* __restrict__ pts,
                                                       uint64_t n, uint4 mlo, uint4 mhi, uint32_t meaning,
                                                       uint8_t* __restrict__ hit, uint32_t* __restrict__ acc)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint32_t a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  bool any = false, outside = false;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t x = quantize(pts[3 * i], side), y = quantize(pts[3 * i + 1], side),
                   z = quantize(pts[3 * i + 2], side);
    uint8_t h = 0;
    if (x < dx && y < dy && z < dz) {
      const uint4* v = reinterpret_cast<const uint4*>(map + (((uint64_t)z * dy + y) * dx + x) * 8);
      const uint4 v0 = v[0], v1 = v[1];
      const uint32_t m[8] = {v0.x & mlo.x, v0.y & mlo.y, v0.z & mlo.z, v0.w & mlo.w,
                             v1.x & mhi.x, v1.y & mhi.y, v1.z & mhi.z, v1.w & mhi.w};
      if (((m[0] & ~1u) | m[1] | m[2] | m[3] | m[4] | m[5] | m[6] | m[7]) != 0u) {  // BitVector::noneButEmpty
#pragma unroll
        for (int k = 0; k < 8; ++k) a[k] |= m[k];
        any = true;
        h = 1;
      }
    } else {
      outside = true;
    }
    hit[i] = h;
  }
  if (any) a[meaning >> 5] |= 1u << (meaning & 31);
#pragma unroll
  for (int k = 0; k < 8; ++k) a[k] = wave_or(a[k]);
  const bool out_w = wave_ballot(outside) != 0;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (a[k]) atomicOr(acc + k, a[k]);
    if (out_w) atomicOr(acc + 8, 1u);
  }
}

// insert pass of the sub-cloud: its meaning into every voxel it reaches,
// eBVM_COLLISION where the test pass flagged the point
__global__ void __launch_bounds__(256) k_selfcoll_insert(uint32_t* __restrict__ map, uint32_t dx, uint32_t dy,
                                                         uint32_t dz, float side, const float* __restrict__ pts,
                                                         uint64_t n, uint32_t meaning, const uint8_t* __restrict__ hit)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t x = quantize(pts[3 * i], side), y = quantize(pts[3 * i + 1], side),
                   z = quantize(pts[3 * i + 2], side);
    if (x < dx && y < dy && z < dz) {
      uint32_t* v = map + (((uint64_t)z * dy + y) * dx + x) * 8;
      atomicOr(v + (meaning >> 5), 1u << (meaning & 31));
      if (hit[i]) atomicOr(v, 1u << M_COLLISION);
    }
  }
}

// kernelClearVoxelMap(bits), VoxelMapOperations.hpp:59-72
__global__ void k_clear_bits(uint32_t* __restrict__ m, uint64_t n, uint4 lo, uint4 hi)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    uint4* p = reinterpret_cast<uint4*>(m + i * 8);
    uint4 a = p[0], b = p[1];
    a.x &= ~lo.x; a.y &= ~lo.y; a.z &= ~lo.z; a.w &= ~lo.w;
    b.x &= ~hi.x; b.y &= ~hi.y; b.z &= ~hi.z; b.w &= ~hi.w;
    p[0] = a;
    p[1] = b;
  }
}

static unsigned stream_grid(uint64_t n)
{
  // enough waves to saturate HBM: ~2048 threads per CU, grid-stride beyond
  uint64_t cap = (uint64_t)(rt().num_cus > 0 ? rt().num_cus : 256) * 8;
  return grid_for(n, 256, (unsigned)cap);
}

int map_clear_async(MapS* m)
{
  if (m->type == TBIT) return fill_u32(m->data, 0u, m->n * 8);
  if (m->type == TPROB) return fill_bytes(m->data, 0x80, m->n);
  return fill_u32(m->data, m->uninit, m->n);
}

// one cloud of a multi-cloud insert: points [ends[c-1], ends[c]) with meanings[c]
template <int T>
__global__ void __launch_bounds__(256) k_insert_cloud(uint8_t* __restrict__ map, uint32_t dx, uint32_t dy, uint32_t dz,
                                                      float side, const float* __restrict__ pts,
                                                      const uint32_t* __restrict__ cloud_ends, uint32_t c,
                                                      const uint32_t* __restrict__ meanings, uint32_t uninit,
                                                      int* __restrict__ outside)
{
  const uint64_t lo = c ? cloud_ends[c - 1] : 0, hi = cloud_ends[c];
  const uint32_t m = meanings ? meanings[c] : 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = lo + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < hi; i += stride) {
    const uint32_t x = quantize(pts[3 * i], side), y = quantize(pts[3 * i + 1], side),
                   z = quantize(pts[3 * i + 2], side);
    const bool in = (x < dx) && (y < dy) && (z < dz);
    if (in) voxel_insert<T>(map, ((uint64_t)z * dy + y) * dx + x, x, y, z, m, uninit);
    else if (outside) *outside = 1;
  }
}

int insert_points_async(MapS* m, const float* pts_dev, uint64_t n, const uint32_t* ends_dev, uint32_t ncl,
                        const uint32_t* meanings_dev, uint32_t meaning, int* outside_dev)
{
  if (n == 0) return GVX_OK;
  unsigned g = grid_for(n, 256, 1u << 16);
  hipStream_t s = rt().stream;
  ProfScope ps("insert_points");
  // Prob and Distance voxels: updates with different meanings do not commute
  // (saturating log-odds of both signs; last-writer distance voxels), so a
  // multi-cloud insert runs cloud by cloud in order = the serial point order.
  // Within a cloud every update is the same and commutes.
  if (m->type != TBIT && ends_dev && meanings_dev && ncl > 1) {
    const unsigned gc = grid_for((n + ncl - 1) / ncl, 256, 1u << 16);
    for (uint32_t c = 0; c < ncl; ++c) {
      if (m->type == TPROB)
        hipLaunchKernelGGL(k_insert_cloud<TPROB>, dim3(gc), dim3(256), 0, s, (uint8_t*)m->data, m->dims[0],
                           m->dims[1], m->dims[2], m->side, pts_dev, ends_dev, c, meanings_dev, m->uninit, outside_dev);
      else
        hipLaunchKernelGGL(k_insert_cloud<TDIST>, dim3(gc), dim3(256), 0, s, (uint8_t*)m->data, m->dims[0],
                           m->dims[1], m->dims[2], m->side, pts_dev, ends_dev, c, meanings_dev, m->uninit, outside_dev);
    }
    GVX_HIP(hipGetLastError());
    return GVX_OK;
  }
#define LAUNCH(T)                                                                                           \
  hipLaunchKernelGGL(k_insert_points<T>, dim3(g), dim3(256), 0, s, (uint8_t*)m->data, m->dims[0], m->dims[1], \
                     m->dims[2], m->side, pts_dev, n, ends_dev, ncl, meanings_dev, meaning, m->uninit, outside_dev)
  if (m->type == TBIT) LAUNCH(TBIT);
  else if (m->type == TPROB) LAUNCH(TPROB);
  else LAUNCH(TDIST);
#undef LAUNCH
  GVX_HIP(hipGetLastError());
  return GVX_OK;
}

// ---------------------------------------------------------------- sensor data
// insertSensorData (ProbVoxelMap.hpp:51-78, kernelInsertSensorData
// VoxelMapOperations.hpp:763-809).  The reference updates voxels with unguarded
// read-modify-writes from all rays at once; here one scan is deterministic: pass 1
// applies the FREE update (-10) to every cell of every ray, pass 2 the OCCUPIED
// update (+72) to every measured voxel.  Updates of one sign commute under the
// saturating rule, so each pass is order independent.

// point -> voxel if it is used: not NaN, point and sensor voxel inside the map,
// and (with a robot map) the robot does not occupy the voxel with robot_bit
__device__ __forceinline__ bool sensor_point(const float* __restrict__ p, float side, uint32_t dx, uint32_t dy,
                                             uint32_t dz, bool sensor_in, const uint8_t* __restrict__ robot,
                                             uint32_t robot_bit, uint32_t& x, uint32_t& y, uint32_t& z)
{
  const float px = p[0], py = p[1], pz = p[2];
  if (!sensor_in || px != px || py != py || pz != pz) return false;
  x = quantize(px, side);
  y = quantize(py, side);
  z = quantize(pz, side);
  if (!(x < dx && y < dy && z < dz)) return false;
  if (robot) {
    const uint8_t b = robot[(((uint64_t)z * dy + y) * dx + x) * 32 + (robot_bit >> 3)];
    if ((b >> (robot_bit & 7)) & 1u) return false;
  }
  return true;
}

// `count` saturating updates by `delta` on the int8 voxel idx (one CAS on its
// word); a saturated voxel is only read
__device__ __forceinline__ void prob_add_n(uint8_t* map, uint64_t idx, int delta, uint32_t count)
{
  uint32_t* w = reinterpret_cast<uint32_t*>(map + (idx & ~(uint64_t)3));
  const uint32_t sh = (uint32_t)(idx & 3) * 8;
  uint32_t old = __atomic_load_n(w, __ATOMIC_RELAXED);
  for (;;) {
    const int cur = (int)(int8_t)((old >> sh) & 0xFF);
    int nv = cur;
    for (uint32_t k = 0; k < count; ++k) {
      const int n2 = prob_update(nv, delta);
      if (n2 == nv) break;
      nv = n2;
    }
    if (nv == cur) return;
    const uint32_t neu = (old & ~(0xFFu << sh)) | ((uint32_t)(uint8_t)(int8_t)nv << sh);
    const uint32_t prev = atomicCAS(w, old, neu);
    if (prev == old) return;
    old = prev;
  }
}

// Points that fall into the same voxel cast the same ray (RayCaster::rayCast walks
// from the sensor voxel to the point's voxel) and add the same OCCUPIED update, so
// the scan runs once per distinct target voxel with its point count as a weight,
// which gives the same result (updates of one sign commute under the saturating
// rule).  The representative of a voxel is its lowest-indexed point, so the
// representatives keep the image order (neighbouring lanes walk neighbouring rays).
//
// tag: hits[voxel] = max over its points of ~i (hits is zero at rest)
__global__ void __launch_bounds__(256) k_sensor_tag(uint32_t* __restrict__ hits, uint32_t dx, uint32_t dy, uint32_t dz,
                                                    float side, bool sensor_in, const float* __restrict__ pts,
                                                    uint64_t n, const uint8_t* __restrict__ robot, uint32_t robot_bit)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    uint32_t x = 0, y = 0, z = 0;
    const bool in = sensor_point(pts + 3 * i, side, dx, dy, dz, sensor_in, robot, robot_bit, x, y, z);
    const uint64_t v = in ? ((uint64_t)z * dy + y) * dx + x : 0ull;
    // a run of consecutive points in one voxel: its first (lowest) index is the run's maximum ~i
    if (wave_run_head(in, v, 0u)) atomicMax(hits + v, ~(uint32_t)i);
  }
}

// rep[i] = 1 when point i is its voxel's representative
__global__ void __launch_bounds__(256) k_sensor_rep(const uint32_t* __restrict__ hits, uint32_t dx, uint32_t dy,
                                                    uint32_t dz, float side, bool sensor_in,
                                                    const float* __restrict__ pts, uint64_t n,
                                                    const uint8_t* __restrict__ robot, uint32_t robot_bit,
                                                    uint32_t* __restrict__ rep)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    uint32_t x, y, z;
    bool r = false;
    if (sensor_point(pts + 3 * i, side, dx, dy, dz, sensor_in, robot, robot_bit, x, y, z))
      r = hits[((uint64_t)z * dy + y) * dx + x] == ~(uint32_t)i;
    rep[i] = r ? 1u : 0u;
  }
}

// weight[pos[r]] += 1 for every point of representative r; list[pos[r]] = voxel of r
__global__ void __launch_bounds__(256) k_sensor_weigh(const uint32_t* __restrict__ hits, uint32_t dx, uint32_t dy,
                                                      uint32_t dz, float side, bool sensor_in,
                                                      const float* __restrict__ pts, uint64_t n,
                                                      const uint8_t* __restrict__ robot, uint32_t robot_bit,
                                                      const uint32_t* __restrict__ pos, uint64_t* __restrict__ list,
                                                      uint32_t* __restrict__ weight)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    uint32_t x = 0, y = 0, z = 0;
    const bool in = sensor_point(pts + 3 * i, side, dx, dy, dz, sensor_in, robot, robot_bit, x, y, z);
    const uint64_t v = in ? ((uint64_t)z * dy + y) * dx + x : 0ull;
    const uint32_t run = wave_run_head(in, v, 0u);  // consecutive points in one voxel add their count once
    if (!in) continue;
    const uint32_t r = ~hits[v];
    if (run) atomicAdd(weight + pos[r], run);
    if (r == (uint32_t)i) list[pos[r]] = v;
  }
}

// hits back to zero under the representatives' voxels
__global__ void __launch_bounds__(256) k_sensor_untag(uint32_t* __restrict__ hits, const uint64_t* __restrict__ list,
                                                      const uint32_t* __restrict__ count)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x, m = *count;
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += stride) hits[list[k]] = 0u;
}

// pass 1: one lane per distinct target voxel walks RayCaster::rayCast's 3-D
// integer line (VoxelMapOperations.h:218-340) from the sensor voxel to the target
// and adds the target's weight to the 32-bit counter of every cell it visits
// (atomics without a return value: the walk never waits).  Rays of neighbouring
// targets (consecutive lanes) share many cells, so at every step each run of
// consecutive lanes on one cell adds its summed weight once (a 640 x 480 depth
// image into a 512^3 map: 21.4 M atomics for 50.7 M cells walked).
__global__ void __launch_bounds__(256) k_sensor_rays(uint32_t* __restrict__ hits, uint32_t dx, uint32_t dy, uint32_t dz,
                                                     uint32_t sx, uint32_t sy, uint32_t sz,
                                                     const uint64_t* __restrict__ list,
                                                     const uint32_t* __restrict__ weight,
                                                     const uint32_t* __restrict__ count)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t first = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t m = *count;
  const uint64_t rounds = (m + stride - 1) / stride;  // uniform trip count: every lane reaches the ballots
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t plane = (uint64_t)dx * dy;
  for (uint64_t r = 0; r < rounds; ++r) {
    const uint64_t k = first + r * stride;
    const bool use = k < m;
    uint32_t wgt = 0;
    int32_t d[3] = {0, 0, 0}, inc[3] = {1, 1, 1};
    if (use) {
      const uint64_t v = list[k];
      wgt = weight[k];
      const uint32_t t[3] = {(uint32_t)(v % dx), (uint32_t)((v / dx) % dy), (uint32_t)(v / plane)};
      const uint32_t f[3] = {sx, sy, sz};
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        if (t[a] > f[a]) { d[a] = (int32_t)(t[a] - f[a]); inc[a] = 1; }
        else { d[a] = (int32_t)(f[a] - t[a]); inc[a] = -1; }
      }
    }
    int64_t left = use ? 1 + (int64_t)d[0] + d[1] + d[2] : 0;
    int32_t exy = d[0] - d[1], exz = d[0] - d[2], eyz = d[1] - d[2];
    const int32_t dx2 = 2 * d[0], dy2 = 2 * d[1], dz2 = 2 * d[2];
    const int64_t sxs = inc[0], sys = inc[1] * (int64_t)dx, szs = inc[2] * (int64_t)plane;
    uint64_t idx = ((uint64_t)sz * dy + sy) * dx + sx;
    for (;;) {
      const bool active = left > 0;
      const uint64_t am = wave_ballot(active);
      if (am == 0) break;
      // consecutive lanes (neighbouring targets) on one cell form a run; the run's
      // first lane adds the run's summed weight with one atomic
      const uint32_t ilo = (uint32_t)idx, ihi = (uint32_t)(idx >> 32);
      const uint32_t plo = __shfl_up(ilo, 1, 64), phi = __shfl_up(ihi, 1, 64);
      const bool same = active && lane > 0 && ((am >> (lane - 1)) & 1ull) != 0ull && plo == ilo && phi == ihi;
      const uint64_t cont = wave_ballot(same);
      uint32_t v = active ? wgt : 0u;
      if (cont != 0ull) {  // uniform: segmented suffix sums of the weights
        const uint64_t above = lane == 63 ? 0ull : ~cont & (~0ull << (lane + 1));  // run starts after this lane
        const uint32_t end = above != 0ull ? (uint32_t)__builtin_ctzll(above) - 1u : 63u;
        for (uint32_t o = 1; o < 64; o <<= 1) {
          const uint32_t t = __shfl_down(v, o, 64);
          if (lane + o <= end) v += t;
        }
      }
      if (active && !same) atomicAdd(hits + idx, v);
      if (active) {
        if (exy > 0 && exz > 0) { idx += sxs; exy -= dy2; exz -= dz2; }
        else if (eyz > 0) { idx += sys; exy += dx2; eyz -= dz2; }
        else { idx += szs; exz += dx2; eyz += dy2; }
        --left;
      }
    }
  }
}

// pass 1b: F free updates in a row on one voxel = max((unknown ? 0 : v) - 10 F,
// -127) (updateOccupancy saturates at MIN_PROBABILITY); the counters go back to
// zero for the next scan.  Four voxels per thread.
__global__ void __launch_bounds__(256) k_sensor_apply(uint8_t* __restrict__ map, uint32_t* __restrict__ hits, uint64_t n)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; 4 * q < n; q += stride) {
    const uint64_t i0 = 4 * q;
    if (i0 + 4 <= n) {
      const uint4 h = reinterpret_cast<const uint4*>(hits)[q];
      if ((h.x | h.y | h.z | h.w) == 0u) continue;
      uint32_t w = reinterpret_cast<const uint32_t*>(map)[q];
      const uint32_t hc[4] = {h.x, h.y, h.z, h.w};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (!hc[k]) continue;
        int v = (int)(int8_t)((w >> (8 * k)) & 0xFFu);
        if (v == PROB_UNKNOWN) v = 0;
        v = max(v - 10 * (int)min(hc[k], 64u), PROB_MIN);
        w = (w & ~(0xFFu << (8 * k))) | ((uint32_t)(uint8_t)(int8_t)v << (8 * k));
      }
      reinterpret_cast<uint32_t*>(map)[q] = w;
      reinterpret_cast<uint4*>(hits)[q] = make_uint4(0u, 0u, 0u, 0u);
    } else {
      for (uint64_t i = i0; i < n; ++i) {
        const uint32_t c = hits[i];
        if (!c) continue;
        int v = (int)(int8_t)map[i];
        if (v == PROB_UNKNOWN) v = 0;
        map[i] = (uint8_t)(int8_t)max(v - 10 * (int)min(c, 64u), PROB_MIN);
        hits[i] = 0u;
      }
    }
  }
}

// pass 2: the measured voxels, each with its point count
__global__ void __launch_bounds__(256) k_sensor_hits(uint8_t* __restrict__ map, const uint64_t* __restrict__ list,
                                                     const uint32_t* __restrict__ weight,
                                                     const uint32_t* __restrict__ count)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x, m = *count;
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += stride)
    prob_add_n(map, list[k], SENSOR_OCC, weight[k]);
}

// ---------------------------------------------------------------- swept volumes
// collisionCheckBitvector with SVCollider(thr, window) for Bit x Bit
// (BitVoxelMap.hpp:90-142, kernelCollideVoxelMapsBitvector
// VoxelMapOperations.hpp:186-239): bitMarginCollisionCheck per voxel, colliding
// voxels of self get eBVM_COLLISION, the mask ORs every colliding voxel's result.
__global__ void __launch_bounds__(256) k_collide_sv(uint8_t* __restrict__ a, const uint8_t* __restrict__ b, uint64_t n,
                                                    uint32_t margin, uint32_t sv_offset,
                                                    unsigned long long* __restrict__ count, uint32_t* __restrict__ mask)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t c = 0;
  uint32_t acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    uint4* pa = reinterpret_cast<uint4*>(a + i * 32);
    const uint4* pb = reinterpret_cast<const uint4*>(b + i * 32);
    const uint4 a0 = pa[0], a1 = pa[1], b0 = pb[0], b1 = pb[1];
    const uint32_t va[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
    const uint32_t vb[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
    bool hit = false;
    // an empty side cannot collide: skip the byte loop for it
    if ((va[0] | va[1] | va[2] | va[3] | va[4] | va[5] | va[6] | va[7]) &&
        ((vb[0] & 0xFFFFFFF0u) | vb[1] | vb[2] | vb[3] | vb[4] | vb[5] | vb[6] | vb[7])) {
      uint32_t r[8];
      hit = bitmargin(va, vb, margin, sv_offset, r);
      if (hit) {
#pragma unroll
        for (int k = 0; k < 8; ++k) acc[k] |= r[k];
        reinterpret_cast<uint32_t*>(pa)[0] = a0.x | (1u << M_COLLISION);
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

// shiftLeftSweptVolumeIDs (BitVoxelMap.hpp:351-362 -> kernelShiftBitVector,
// VoxelMapOperations.hpp:815-823): performLeftShift on every 32-byte voxel
__global__ void __launch_bounds__(256) k_shift_sv(uint8_t* __restrict__ v, uint64_t n, uint32_t shift)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    uint4* p = reinterpret_cast<uint4*>(v + i * 32);
    const uint4 a0 = p[0], a1 = p[1];
    uint32_t w[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
    left_shift_sv(w, shift);
    p[0] = make_uint4(w[0], w[1], w[2], w[3]);
    p[1] = make_uint4(w[4], w[5], w[6], w[7]);
  }
}

// per-voxel ray counters of insertSensorData: allocated zeroed once per size and
// left zeroed by every scan (k_sensor_apply clears what it consumed)
static int sensor_hits_buffer(uint64_t n, uint32_t** out)
{
  static uint32_t* buf = nullptr;
  static uint64_t cap = 0;
  if (cap < n) {
    GVX_CHECK(sync());
    dfree(buf);
    buf = nullptr;
    cap = 0;
    GVX_CHECK(dalloc((void**)&buf, ((n + 3) / 4) * 16));
    GVX_CHECK(fill_u32(buf, 0u, ((n + 3) / 4) * 4));
    cap = n;
  }
  *out = buf;
  return GVX_OK;
}

// per-point arrays of insertSensorData (grown on demand, kept): representative flags
// and their scan (n + 1 words), weights (n words), target voxels (n u64)
static int sensor_scratch(uint64_t n, uint32_t** rep, uint32_t** pos, uint32_t** weight, uint64_t** list)
{
  static uint8_t* buf = nullptr;
  static uint64_t cap = 0;
  const auto al = [](uint64_t b) { return (b + 255) & ~(uint64_t)255; };
  if (cap < n) {
    GVX_CHECK(sync());
    dfree(buf);
    buf = nullptr;
    cap = 0;
    GVX_CHECK(dalloc((void**)&buf, al(4 * n) + al(4 * (n + 1)) + al(4 * n) + al(8 * n)));
    cap = n;
  }
  *rep = (uint32_t*)buf;
  *pos = (uint32_t*)(buf + al(4 * cap));
  *weight = (uint32_t*)(buf + al(4 * cap) + al(4 * (cap + 1)));
  *list = (uint64_t*)(buf + al(4 * cap) + al(4 * (cap + 1)) + al(4 * cap));
  return GVX_OK;
}

// per-point flags of the self-collision test (grown on demand, kept)
static int selfcoll_flags(uint64_t n, uint8_t** out)
{
  static uint8_t* buf = nullptr;
  static uint64_t cap = 0;
  if (cap < n) {
    GVX_CHECK(sync());
    dfree(buf);
    buf = nullptr;
    cap = 0;
    GVX_CHECK(dalloc((void**)&buf, (n + 255) & ~(uint64_t)255));
    cap = n;
  }
  *out = buf;
  return GVX_OK;
}

// insertMetaPointCloudWithSelfCollisionCheck on device points: sub-clouds in
// order, each a test launch and an insert launch.  acc_dev: 9 words (colliding
// meanings, points-outside flag), cleared here.  meanings / masks on the host,
// NULL for the defaults (BitVoxelMap.hpp:262-292).
int selfcoll_insert_async(MapS* m, const float* pts_dev, const uint32_t* sizes, uint32_t ncl,
                          const uint32_t* meanings, const uint8_t* masks, uint32_t* acc_dev)
{
  if (m->type != TBIT)
    return fail(GVX_ERR_UNSUPPORTED, "insertMetaPointCloudWithSelfCollisionCheck: BitVector maps only");
  uint64_t total = 0;
  for (uint32_t c = 0; c < ncl; ++c) {
    total += sizes[c];
    const uint32_t mc = meanings ? meanings[c] : M_SV_START + c;
    if (mc > 255) return fail(GVX_ERR_BAD_ARG, "meaning > 255");
  }
  uint8_t* hit = nullptr;
  GVX_CHECK(selfcoll_flags(total ? total : 1, &hit));
  GVX_CHECK(fill_u32(acc_dev, 0u, 12));
  hipStream_t s = rt().stream;
  uint64_t off = 0;
  for (uint32_t c = 0; c < ncl; ++c) {
    const uint32_t mc = meanings ? meanings[c] : M_SV_START + c;
    uint32_t w[8];
    if (masks) {
      memcpy(w, masks + (uint64_t)c * 32, 32);
    } else {
      for (int k = 0; k < 8; ++k) w[k] = 0xFFFFFFFFu;
      w[M_COLLISION >> 5] &= ~(1u << (M_COLLISION & 31));
      w[mc >> 5] &= ~(1u << (mc & 31));
    }
    const uint64_t n = sizes[c];
    if (n) {
      const unsigned g = grid_for(n, 256, 1u << 14);
      hipLaunchKernelGGL(k_selfcoll_test, dim3(g), dim3(256), 0, s, (const uint32_t*)m->data, m->dims[0], m->dims[1],
                         m->dims[2], m->side, pts_dev + 3 * off, n, make_uint4(w[0], w[1], w[2], w[3]),
                         make_uint4(w[4], w[5], w[6], w[7]), mc, hit + off, acc_dev);
      hipLaunchKernelGGL(k_selfcoll_insert, dim3(g), dim3(256), 0, s, (uint32_t*)m->data, m->dims[0], m->dims[1],
                         m->dims[2], m->side, pts_dev + 3 * off, n, mc, (const uint8_t*)(hit + off));
      GVX_HIP(hipGetLastError());
    }
    off += n;
  }
  return GVX_OK;
}

// colliding meanings and flags back from acc (after selfcoll_insert_async)
int selfcoll_finish(const uint32_t* acc_dev, uint8_t colliding[32], int* self_collision, int* points_outside)
{
  Runtime& r = rt();
  GVX_HIP(hipMemcpyAsync(r.host_small, acc_dev, 36, hipMemcpyDeviceToHost, r.stream));
  GVX_CHECK(sync());
  const uint32_t* h = (const uint32_t*)r.host_small;
  uint32_t any = 0;
  for (int k = 0; k < 8; ++k) any |= h[k];
  if (colliding) memcpy(colliding, h, 32);
  if (self_collision) *self_collision = any != 0u;
  if (points_outside) *points_outside = (int)h[8];
  return GVX_OK;
}

int shift_sv_async(uint8_t* voxels, uint64_t n, uint32_t shift)
{
  if (!n) return GVX_OK;
  hipLaunchKernelGGL(k_shift_sv, dim3(grid_for(n, 256, 1u << 16)), dim3(256), 0, rt().stream, voxels, n, shift);
  GVX_HIP(hipGetLastError());
  return GVX_OK;
}

}  // namespace gvx

using namespace gvx;

static int check_map(gvx_map m)
{
  if (!m || !m->data) return fail(GVX_ERR_BAD_ARG, "null map handle");
  return GVX_OK;
}

extern "C" {

int gvx_map_create(int type, uint32_t dx, uint32_t dy, uint32_t dz, float side, gvx_map* out)
{
  if (!out) return fail(GVX_ERR_BAD_ARG, "gvx_map_create: out is null");
  *out = nullptr;
  GVX_CHECK(ensure_init());
  if (type != TBIT && type != TPROB && type != TDIST)
    return fail(GVX_ERR_UNSUPPORTED, "THIS TYPE OF DATA STRUCTURE IS NOT YET IMPLEMENTED!");
  if (dx == 0 || dy == 0 || dz == 0 || !(side > 0.0f)) return fail(GVX_ERR_BAD_ARG, "gvx_map_create: bad dims/side");
  if (type == TDIST && (dx > 1024 || dy > 1024 || dz > 4094))
    return fail(GVX_ERR_UNSUPPORTED, "DistanceVoxelMap: x,y <= 1024 and z <= 4094 (10/10/12-bit packing)");
  uint64_t n = (uint64_t)dx * dy * dz;
  if (n > (1ull << 32)) return fail(GVX_ERR_UNSUPPORTED, "map larger than 2^32 voxels");
  gvx_map m = new gvx_map_s();
  m->type = type;
  m->dims[0] = dx; m->dims[1] = dy; m->dims[2] = dz;
  m->side = side;
  m->n = n;
  m->bytes = n * voxel_bytes(type);
  m->uninit = dist_uninit(m->dims);
  size_t alloc = (m->bytes + 255) & ~(size_t)255;
  int s = dalloc(&m->data, alloc);
  if (s != GVX_OK) { delete m; return s; }
  s = map_clear_async(m);
  if (s == GVX_OK) s = sync();
  if (s != GVX_OK) { dfree(m->data); delete m; return s; }
  *out = m;
  return GVX_OK;
}

int gvx_map_destroy(gvx_map m)
{
  if (!m) return GVX_OK;
  (void)sync();
  dfree(m->data);
  delete m;
  return GVX_OK;
}

int gvx_map_info(gvx_map m, int* type, uint32_t dims[3], float* side, uint64_t* bytes)
{
  GVX_CHECK(check_map(m));
  if (type) *type = m->type;
  if (dims) { dims[0] = m->dims[0]; dims[1] = m->dims[1]; dims[2] = m->dims[2]; }
  if (side) *side = m->side;
  if (bytes) *bytes = m->bytes;
  return GVX_OK;
}

int gvx_map_device_ptr(gvx_map m, void** ptr)
{
  GVX_CHECK(check_map(m));
  if (!ptr) return fail(GVX_ERR_BAD_ARG, "null");
  *ptr = m->data;
  return GVX_OK;
}

int gvx_map_clear(gvx_map m)
{
  GVX_CHECK(check_map(m));
  std::lock_guard<std::recursive_mutex> g(m->mu);
  {
    ProfScope ps("clear_map");
    GVX_CHECK(map_clear_async(m));
  }
  return sync();
}

int gvx_map_clear_bits(gvx_map m, const uint8_t bits[32])
{
  GVX_CHECK(check_map(m));
  if (m->type != TBIT) return fail(GVX_ERR_UNSUPPORTED, "clearBits needs a BitVector map");
  if (!bits) return fail(GVX_ERR_BAD_ARG, "null bits");
  std::lock_guard<std::recursive_mutex> g(m->mu);
  uint32_t w[8];
  memcpy(w, bits, 32);
  hipLaunchKernelGGL(k_clear_bits, dim3(stream_grid(m->n)), dim3(256), 0, rt().stream, (uint32_t*)m->data, m->n,
                     make_uint4(w[0], w[1], w[2], w[3]), make_uint4(w[4], w[5], w[6], w[7]));
  GVX_HIP(hipGetLastError());
  return sync();
}

int gvx_map_read(gvx_map m, void* dst, int mem)
{
  GVX_CHECK(check_map(m));
  if (!dst) return fail(GVX_ERR_BAD_ARG, "null dst");
  std::lock_guard<std::recursive_mutex> g(m->mu);
  GVX_CHECK(copy_out(dst, m->data, m->bytes, mem));
  return sync();
}

int gvx_map_write(gvx_map m, const void* src, int mem)
{
  GVX_CHECK(check_map(m));
  if (!src) return fail(GVX_ERR_BAD_ARG, "null src");
  std::lock_guard<std::recursive_mutex> g(m->mu);
  GVX_CHECK(copy_in(m->data, src, m->bytes, mem));
  return sync();
}

static int finish_outside(int* points_outside)
{
  Runtime& r = rt();
  if (points_outside) {
    GVX_HIP(hipMemcpyAsync(r.host_small, r.dev_small, sizeof(int), hipMemcpyDeviceToHost, r.stream));
  }
  GVX_CHECK(sync());
  if (points_outside) *points_outside = *(int*)r.host_small;
  return GVX_OK;
}

int gvx_map_insert_points(gvx_map m, const float* xyz, uint64_t n, uint32_t meaning, int mem, int* points_outside)
{
  GVX_CHECK(check_map(m));
  if (n && !xyz) return fail(GVX_ERR_BAD_ARG, "null points");
  if (meaning > 255) return fail(GVX_ERR_BAD_ARG, "meaning > 255");
  std::lock_guard<std::recursive_mutex> g(m->mu);
  std::lock_guard<std::recursive_mutex> g2(rt().mu);
  const void* d = nullptr;
  GVX_CHECK(stage_input(xyz, n * 12, mem, &d));
  GVX_CHECK(fill_u32(rt().dev_small, 0u, 4));
  GVX_CHECK(insert_points_async(m, (const float*)d, n, nullptr, 0, nullptr, meaning, (int*)rt().dev_small));
  return finish_outside(points_outside);
}

int gvx_map_insert_sensor(gvx_map m, const float* xyz, uint64_t n, const float sensor_pose[3], int enable_raycasting,
                          const void* robot_bits_dev, uint32_t robot_meaning, int mem)
{
  GVX_CHECK(check_map(m));
  if (m->type != TPROB) return fail(GVX_ERR_UNSUPPORTED, "insertSensorData needs a ProbVoxelMap");
  if (!sensor_pose) return fail(GVX_ERR_BAD_ARG, "null sensor pose");
  if (n && !xyz) return fail(GVX_ERR_BAD_ARG, "null points");
  if (robot_meaning > 255) return fail(GVX_ERR_BAD_ARG, "meaning > 255");
  if (n > m->n) n = m->n;  // the reference's loop also stops at the map size
  std::lock_guard<std::recursive_mutex> g(m->mu);
  std::lock_guard<std::recursive_mutex> g2(rt().mu);
  if (!n) return GVX_OK;
  const void* d = nullptr;
  GVX_CHECK(stage_input(xyz, n * 12, mem, &d));
  const float side = m->side;
  const uint32_t sx = quantize(sensor_pose[0], side), sy = quantize(sensor_pose[1], side),
                 sz = quantize(sensor_pose[2], side);
  const bool sin = sx < m->dims[0] && sy < m->dims[1] && sz < m->dims[2];
  const uint8_t* robot = (const uint8_t*)robot_bits_dev;
  hipStream_t s = rt().stream;
  if (n >= 0xFFFFFFFFull) return fail(GVX_ERR_UNSUPPORTED, "insertSensorData: more than 2^32 - 2 points");
  uint32_t* hits = nullptr;
  GVX_CHECK(sensor_hits_buffer(m->n, &hits));
  uint32_t *rep = nullptr, *pos = nullptr, *weight = nullptr;
  uint64_t* list = nullptr;
  GVX_CHECK(sensor_scratch(n, &rep, &pos, &weight, &list));
  const uint32_t* count = pos + n;
  const unsigned gp = grid_for(n, 256, 1u << 14);
  {
    // distinct target voxels (in image order) and their point counts
    ProfScope ps("sensor_targets");
    const float* pts = (const float*)d;
    hipLaunchKernelGGL(k_sensor_tag, dim3(gp), dim3(256), 0, s, hits, m->dims[0], m->dims[1], m->dims[2], side, sin,
                       pts, n, robot, robot_meaning);
    hipLaunchKernelGGL(k_sensor_rep, dim3(gp), dim3(256), 0, s, hits, m->dims[0], m->dims[1], m->dims[2], side, sin,
                       pts, n, robot, robot_meaning, rep);
    GVX_HIP(hipGetLastError());
    GVX_CHECK(scan_u32(rep, pos, n, pos + n));
    GVX_CHECK(fill_u32(weight, 0u, n));
    hipLaunchKernelGGL(k_sensor_weigh, dim3(gp), dim3(256), 0, s, hits, m->dims[0], m->dims[1], m->dims[2], side,
                       sin, pts, n, robot, robot_meaning, pos, list, weight);
    hipLaunchKernelGGL(k_sensor_untag, dim3(gp), dim3(256), 0, s, hits, list, count);
    GVX_HIP(hipGetLastError());
  }
  if (enable_raycasting) {
    {
      ProfScope ps("sensor_rays");
      hipLaunchKernelGGL(k_sensor_rays, dim3(gp), dim3(256), 0, s, hits, m->dims[0], m->dims[1], m->dims[2], sx, sy,
                         sz, list, weight, count);
      GVX_HIP(hipGetLastError());
    }
    {
      ProfScope ps("sensor_apply");
      hipLaunchKernelGGL(k_sensor_apply, dim3(grid_for((m->n + 3) / 4, 256, 1u << 16)), dim3(256), 0, s,
                         (uint8_t*)m->data, hits, m->n);
      GVX_HIP(hipGetLastError());
    }
  }
  {
    ProfScope ps("sensor_hits");
    hipLaunchKernelGGL(k_sensor_hits, dim3(grid_for(n, 256, 1u << 16)), dim3(256), 0, s, (uint8_t*)m->data, list,
                       weight, count);
    GVX_HIP(hipGetLastError());
  }
  return sync();
}

int gvx_map_insert_coords(gvx_map m, const uint32_t* xyz, uint64_t n, uint32_t meaning, int mem, int* points_outside)
{
  GVX_CHECK(check_map(m));
  if (n && !xyz) return fail(GVX_ERR_BAD_ARG, "null coords");
  if (meaning > 255) return fail(GVX_ERR_BAD_ARG, "meaning > 255");
  std::lock_guard<std::recursive_mutex> g(m->mu);
  std::lock_guard<std::recursive_mutex> g2(rt().mu);
  const void* d = nullptr;
  GVX_CHECK(stage_input(xyz, n * 12, mem, &d));
  GVX_CHECK(fill_u32(rt().dev_small, 0u, 4));
  if (n) {
    unsigned gr = grid_for(n, 256, 1u << 16);
    hipStream_t s = rt().stream;
#define LAUNCH(T)                                                                                               \
  hipLaunchKernelGGL(k_insert_coords<T>, dim3(gr), dim3(256), 0, s, (uint8_t*)m->data, m->dims[0], m->dims[1], \
                     m->dims[2], (const uint32_t*)d, n, meaning, m->uninit, (int*)rt().dev_small)
    if (m->type == TBIT) LAUNCH(TBIT);
    else if (m->type == TPROB) LAUNCH(TPROB);
    else LAUNCH(TDIST);
#undef LAUNCH
    GVX_HIP(hipGetLastError());
  }
  return finish_outside(points_outside);
}

int gvx_map_insert_meta(gvx_map m, const float* xyz, const uint32_t* sizes, uint32_t ncl, const uint32_t* meanings,
                        uint32_t meaning, int mem, int* points_outside)
{
  GVX_CHECK(check_map(m));
  if (ncl == 0) return finish_outside(points_outside);
  if (!sizes) return fail(GVX_ERR_BAD_ARG, "null cloud sizes");
  std::vector<uint32_t> tab(2 * (size_t)ncl);
  uint64_t total = 0;
  for (uint32_t c = 0; c < ncl; ++c) {
    total += sizes[c];
    if (total > 0xFFFFFFFFull) return fail(GVX_ERR_UNSUPPORTED, "meta cloud larger than 2^32 points");
    tab[c] = (uint32_t)total;
    uint32_t mc = meanings ? meanings[c] : meaning;
    if (mc > 255) return fail(GVX_ERR_BAD_ARG, "meaning > 255");
    tab[ncl + c] = mc;
  }
  if (total && !xyz) return fail(GVX_ERR_BAD_ARG, "null points");
  std::lock_guard<std::recursive_mutex> g(m->mu);
  std::lock_guard<std::recursive_mutex> g2(rt().mu);
  // staging: [points | table]
  size_t pb = (total * 12 + 255) & ~(size_t)255;
  void* st = nullptr;
  GVX_CHECK(stage_reserve(pb + tab.size() * 4, &st));
  const float* dp = (const float*)xyz;
  if (mem == GVX_MEM_HOST) {
    GVX_HIP(hipMemcpyAsync(st, xyz, total * 12, hipMemcpyHostToDevice, rt().stream));
    dp = (const float*)st;
  }
  uint32_t* dtab = (uint32_t*)((char*)st + pb);
  GVX_HIP(hipMemcpyAsync(dtab, tab.data(), tab.size() * 4, hipMemcpyHostToDevice, rt().stream));
  GVX_CHECK(fill_u32(rt().dev_small, 0u, 4));
  GVX_CHECK(insert_points_async(m, dp, total, dtab, ncl, dtab + ncl, meaning, (int*)rt().dev_small));
  return finish_outside(points_outside);
}

int gvx_map_insert_meta_selfcoll(gvx_map m, const float* xyz, const uint32_t* cloud_sizes, uint32_t n_clouds,
                                 const uint32_t* meanings, const uint8_t* masks, int mem, uint8_t colliding[32],
                                 int* self_collision, int* points_outside)
{
  GVX_CHECK(check_map(m));
  if (n_clouds && !cloud_sizes) return fail(GVX_ERR_BAD_ARG, "null cloud sizes");
  uint64_t total = 0;
  for (uint32_t c = 0; c < n_clouds; ++c) total += cloud_sizes[c];
  if (total && !xyz) return fail(GVX_ERR_BAD_ARG, "null points");
  std::lock_guard<std::recursive_mutex> g(m->mu);
  std::lock_guard<std::recursive_mutex> g2(rt().mu);
  const void* d = nullptr;
  GVX_CHECK(stage_input(xyz, total * 12, mem, &d));
  uint32_t* acc = (uint32_t*)rt().dev_small + 64;
  GVX_CHECK(selfcoll_insert_async(m, (const float*)d, cloud_sizes, n_clouds, meanings, masks, acc));
  return selfcoll_finish(acc, colliding, self_collision, points_outside);
}

int gvx_map_collide(gvx_map self, gvx_map other, float threshold, const int32_t offset[3], uint64_t* count)
{
  GVX_CHECK(check_map(self));
  GVX_CHECK(check_map(other));
  if (!count) return fail(GVX_ERR_BAD_ARG, "null count");
  if (memcmp(self->dims, other->dims, sizeof(self->dims)) != 0)
    return fail(GVX_ERR_DIM_MISMATCH, "collideWith: map dimensions differ");
  if (self->type == TDIST || other->type == TDIST) {  // DefaultCollider: distance voxels never collide
    *count = 0;
    return GVX_OK;
  }
  std::unique_lock<std::recursive_mutex> g1(self->mu, std::defer_lock), g2(other->mu, std::defer_lock);
  if (self == other) g1.lock(); else std::lock(g1, g2);
  std::lock_guard<std::recursive_mutex> g3(rt().mu);
  const uint64_t n = self->n;
  int64_t off = 0;
  if (offset)
    off = (int64_t)offset[2] * (int64_t)self->dims[0] * (int64_t)self->dims[1] +
          (int64_t)offset[1] * (int64_t)self->dims[0] + (int64_t)offset[0];
  uint64_t lo = off < 0 ? (uint64_t)(-off) : 0, hi = off > 0 ? (off >= (int64_t)n ? 0 : n - (uint64_t)off) : n;
  if (lo > n) lo = n;
  const int thr = float_to_probability(threshold);
  unsigned long long* dcount = (unsigned long long*)rt().dev_small;
  GVX_CHECK(fill_u32(dcount, 0u, 4));
  if (hi > lo) {
    unsigned g = stream_grid(hi - lo);
    hipStream_t s = rt().stream;
    uint8_t* a = (uint8_t*)self->data;
    const uint8_t* b = (const uint8_t*)other->data;
    ProfScope ps("collide");
    const unsigned g16 = stream_grid((hi - lo + 15) / 16 + 1);
    if (self->type == TBIT && other->type == TBIT)
      hipLaunchKernelGGL((k_collide<TBIT, TBIT>), dim3(g), dim3(256), 0, s, a, b, lo, hi, off, thr, dcount);
    else if (self->type == TBIT && other->type == TPROB)
      hipLaunchKernelGGL((k_collide_prob<TBIT, TPROB, false>), dim3(g16), dim3(256), 0, s, a, b, lo, hi, off, thr,
                         dcount);
    else if (self->type == TPROB && other->type == TBIT)
      hipLaunchKernelGGL((k_collide_prob<TPROB, TBIT, true>), dim3(g16), dim3(256), 0, s, a, b, lo, hi, off, thr,
                         dcount);
    else
      hipLaunchKernelGGL((k_collide_prob<TPROB, TPROB, false>), dim3(g16), dim3(256), 0, s, a, b, lo, hi, off, thr,
                         dcount);
    GVX_HIP(hipGetLastError());
  }
  GVX_HIP(hipMemcpyAsync(rt().host_small, dcount, 8, hipMemcpyDeviceToHost, rt().stream));
  GVX_CHECK(sync());
  *count = *(unsigned long long*)rt().host_small;
  return GVX_OK;
}

int gvx_map_collide_types(gvx_map self, gvx_map other, float threshold, uint64_t* count, uint8_t mask[32])
{
  GVX_CHECK(check_map(self));
  GVX_CHECK(check_map(other));
  if (!count) return fail(GVX_ERR_BAD_ARG, "null count");
  if (self->type != TBIT) return fail(GVX_ERR_UNSUPPORTED, "collideWithTypes: self must be a BitVector map");
  if (other->type != TBIT && other->type != TPROB)
    return fail(GVX_ERR_UNSUPPORTED, "collideWithTypes: other must be a BitVector or Prob map");
  if (memcmp(self->dims, other->dims, sizeof(self->dims)) != 0)
    return fail(GVX_ERR_DIM_MISMATCH, "collideWithTypes: map dimensions differ");
  std::unique_lock<std::recursive_mutex> g1(self->mu, std::defer_lock), g2(other->mu, std::defer_lock);
  if (self == other) g1.lock(); else std::lock(g1, g2);
  std::lock_guard<std::recursive_mutex> g3(rt().mu);
  const int thr = sv_float_to_probability(threshold);
  unsigned long long* dcount = (unsigned long long*)rt().dev_small;
  uint32_t* dmask = (uint32_t*)rt().dev_small + 2;
  GVX_CHECK(fill_u32(rt().dev_small, 0u, 12));
  {
    unsigned g = stream_grid(self->n);
    ProfScope ps("collide_types");
    if (other->type == TBIT)
      hipLaunchKernelGGL(k_collide_types, dim3(g), dim3(256), 0, rt().stream, (uint8_t*)self->data,
                         (const uint8_t*)other->data, self->n, dcount, dmask);
    else
      hipLaunchKernelGGL(k_collide_types_prob, dim3(stream_grid((self->n + 15) / 16)), dim3(256), 0, rt().stream,
                         (uint8_t*)self->data, (const uint8_t*)other->data, self->n, thr, dcount, dmask);
    GVX_HIP(hipGetLastError());
  }
  GVX_HIP(hipMemcpyAsync(rt().host_small, rt().dev_small, 40, hipMemcpyDeviceToHost, rt().stream));
  GVX_CHECK(sync());
  *count = *(unsigned long long*)rt().host_small;
  if (mask) memcpy(mask, (uint8_t*)rt().host_small + 8, 32);
  return GVX_OK;
}

int gvx_map_collide_sv(gvx_map self, gvx_map other, float threshold, uint32_t window, uint32_t sv_offset,
                       uint64_t* count, uint8_t mask[32])
{
  GVX_CHECK(check_map(self));
  GVX_CHECK(check_map(other));
  if (!count) return fail(GVX_ERR_BAD_ARG, "null count");
  if (self->type != TBIT) return fail(GVX_ERR_UNSUPPORTED, "collisionCheckBitvector: self must be a BitVector map");
  if (other->type == TPROB) return gvx_map_collide_types(self, other, threshold, count, mask);  // window unused
  if (other->type != TBIT) return fail(GVX_ERR_UNSUPPORTED, "collisionCheckBitvector: other must be BitVector or Prob");
  if (window > 31 || sv_offset / 8 > 31 - window)
    return fail(GVX_ERR_BAD_ARG, "SV window must be <= 31 and sv_offset / 8 <= 31 - window");
  if (memcmp(self->dims, other->dims, sizeof(self->dims)) != 0)
    return fail(GVX_ERR_DIM_MISMATCH, "collisionCheckBitvector: map dimensions differ");
  std::unique_lock<std::recursive_mutex> g1(self->mu, std::defer_lock), g2(other->mu, std::defer_lock);
  if (self == other) g1.lock(); else std::lock(g1, g2);
  std::lock_guard<std::recursive_mutex> g3(rt().mu);
  unsigned long long* dcount = (unsigned long long*)rt().dev_small;
  uint32_t* dmask = (uint32_t*)rt().dev_small + 2;
  GVX_CHECK(fill_u32(rt().dev_small, 0u, 12));
  {
    ProfScope ps("collide_sv");
    hipLaunchKernelGGL(k_collide_sv, dim3(stream_grid(self->n)), dim3(256), 0, rt().stream, (uint8_t*)self->data,
                       (const uint8_t*)other->data, self->n, window, sv_offset, dcount, dmask);
    GVX_HIP(hipGetLastError());
  }
  GVX_HIP(hipMemcpyAsync(rt().host_small, rt().dev_small, 40, hipMemcpyDeviceToHost, rt().stream));
  GVX_CHECK(sync());
  *count = *(unsigned long long*)rt().host_small;
  if (mask) memcpy(mask, (uint8_t*)rt().host_small + 8, 32);
  return GVX_OK;
}

int gvx_map_shift_sv(gvx_map m, uint32_t shift)
{
  GVX_CHECK(check_map(m));
  if (m->type != TBIT) return fail(GVX_ERR_UNSUPPORTED, "shiftLeftSweptVolumeIDs needs a BitVector map");
  if (shift > 56) return fail(GVX_ERR_BAD_ARG, "Maximum shift size is 56");
  std::lock_guard<std::recursive_mutex> g(m->mu);
  ProfScope ps("shift_sv");
  GVX_CHECK(shift_sv_async((uint8_t*)m->data, m->n, shift));
  return sync();
}

}  // extern "C"
