// gvx_internal.hpp -- shared runtime state and device helpers of libgvx (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gvx.h"

namespace gvx {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define GVX_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return ::gvx::fail(GVX_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define GVX_CHECK(call)       \
  do {                        \
    int s_ = (call);          \
    if (s_ != GVX_OK) return s_; \
  } while (0)

// ---------------------------------------------------------------- runtime
struct Runtime {
  int device = -1;
  hipStream_t stream = nullptr;
  int num_cus = 0;
  std::recursive_mutex mu;
  // device staging buffer for host inputs, grown on demand
  void* stage = nullptr;
  size_t stage_bytes = 0;
  // small pinned host + device scratch for results
  void* host_small = nullptr;
  void* dev_small = nullptr;
  bool profiling = false;
};
Runtime& rt();
int ensure_init();
// takes the device address of the EDT workspace of gvx_dist_pba_noalloc (at gvx_init)
int pba_noalloc_setup(int device);
// device copy of a caller buffer (host -> staging, device -> as is)
int stage_input(const void* src, size_t bytes, int mem, const void** dev_out);
int stage_reserve(size_t bytes, void** dev_out);
int sync();
int dalloc(void** p, size_t bytes);
void dfree(void* p);
int copy_out(void* dst, const void* dev_src, size_t bytes, int mem);
int copy_in(void* dev_dst, const void* src, size_t bytes, int mem);

// named-kernel HIP-event timing on the library stream
struct ProfScope {
  const char* name;
  hipEvent_t a = nullptr, b = nullptr;
  explicit ProfScope(const char* n);
  ~ProfScope();
};

// fill kernels (used instead of hipMemsetAsync so that work stays graph-capturable)
int fill_u32(void* dst, uint32_t value, uint64_t n_words);
int fill_bytes(void* dst, uint8_t value, uint64_t n_bytes);

// exclusive scan of uint32 counts into uint32 offsets (n entries), total in *total_dev (may be null)
int scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* total_dev);

inline unsigned grid_for(uint64_t n, unsigned block, unsigned cap = 1u << 20)
{
  uint64_t g = (n + block - 1) / block;
  if (g == 0) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

// ---------------------------------------------------------------- constants
constexpr uint32_t M_FREE = 0, M_OCCUPIED = 1, M_COLLISION = 2, M_UNKNOWN = 3;
constexpr uint32_t M_SV_START = 4, M_SV_END = 254, M_UNCERTAIN = 129, M_UNDEFINED = 255;
constexpr int PROB_MIN = -127, PROB_MAX = 127;
constexpr int8_t PROB_UNKNOWN = -128;
constexpr int SENSOR_FREE = -10, SENSOR_OCC = 72;

inline uint32_t dist_uninit(const uint32_t d[3])
{
  return (d[0] <= 1023 && d[1] <= 1023 && d[2] <= 1023) ? 0x3FFFFFFFu : 0xFFFFFFFFu;
}

// host-side threshold conversions (evaluated on the CPU, no fused multiply-add)
int8_t float_to_probability(float v);     // ProbabilisticVoxel.h:39-45 (clamped)
int8_t sv_float_to_probability(float v);  // SVCollider.hpp:133-137 (not clamped)

// ---------------------------------------------------------------- handles
struct MapS {
  int type;
  uint32_t dims[3];
  float side;
  uint64_t n;       // voxels
  uint64_t bytes;   // allocation
  void* data = nullptr;
  uint32_t uninit = 0;
  std::recursive_mutex mu;
};

inline uint64_t voxel_bytes(int type)
{
  return type == GVX_MT_BITVECTOR_VOXELMAP ? 32 : (type == GVX_MT_PROBAB_VOXELMAP ? 1 : 4);
}

// ---------------------------------------------------------------- device helpers
// floor(p / side) saturated to uint32 (VoxelMapOperations.h:121-131; the GPU's
// float->uint conversion clamps, NaN -> 0).  IEEE division.  Branch-free: the
// clamp (fmaxf drops NaN) leaves a value the conversion takes as is; quotients
// of 2^32 - 256 and more give 2^32 - 256, which every caller treats as outside
// the map as it would 2^32 - 1.
__host__ __device__ __forceinline__ uint32_t quantize(float p, float side)
{
  const float f = fminf(fmaxf(floorf(p / side), 0.0f), 4294967040.0f);
  return (uint32_t)f;
}

__device__ __forceinline__ int32_t quantize_signed(float p, float side)
{
  float f = floorf(p / side);
  if (f != f) return 0;
  if (f <= -2147483648.0f) return INT32_MIN;
  if (f >= 2147483648.0f) return INT32_MAX;
  return (int32_t)f;
}

// p' = M p with M a row-major 3x4 affine: fmaf(a13, z, fmaf(a12, y, fmaf(a11, x, a14)))
__device__ __forceinline__ void transform(const float* m, float x, float y, float z, float& ox, float& oy,
                                          float& oz)
{
  ox = fmaf(m[2], z, fmaf(m[1], y, fmaf(m[0], x, m[3])));
  oy = fmaf(m[6], z, fmaf(m[5], y, fmaf(m[4], x, m[7])));
  oz = fmaf(m[10], z, fmaf(m[9], y, fmaf(m[8], x, m[11])));
}

// ProbabilisticVoxel::updateOccupancy / insert (voxel/ProbabilisticVoxel.hpp:43-101)
__host__ __device__ __forceinline__ int prob_update(int cur, int delta)
{
  if (cur == PROB_UNKNOWN) return delta;
  int v = cur + delta;
  v = v < PROB_MIN ? PROB_MIN : v;
  v = v > PROB_MAX ? PROB_MAX : v;
  return v;
}
__host__ __device__ __forceinline__ int prob_insert(int cur, uint32_t meaning)
{
  switch (meaning) {
    case M_FREE: return prob_update(cur, SENSOR_FREE);
    case M_COLLISION:
    case M_OCCUPIED: return prob_update(cur, SENSOR_OCC);
    case M_UNKNOWN:
    case M_UNDEFINED: return PROB_UNKNOWN;
    default:
      if (meaning <= M_SV_END && meaning >= M_SV_START) return prob_update(cur, (int)meaning - (int)M_UNCERTAIN);
      return PROB_UNKNOWN;
  }
}

// BitVector::noneButEmpty on the eight 32-bit words of a voxel (BitVector.h:183-198)
__device__ __forceinline__ bool bits_occupied(const uint4& a, const uint4& b)
{
  return ((a.x & ~1u) | a.y | a.z | a.w | b.x | b.y | b.z | b.w) != 0u;
}

// byte k of a 256-bit vector held as 8 words (bit i at byte i >> 3, as the voxel layout)
__device__ __forceinline__ uint32_t bv_byte(const uint32_t v[8], uint32_t k) { return (v[k >> 2] >> ((k & 3) * 8)) & 0xFFu; }

// bitMarginCollisionCheck (helpers/BitVector.h:415-471), the same steps on a
// register copy: a 64-bit buffer slides over v2's bytes (its byte 0 without bits
// 0..3); byte k of the result ORs, over the 2 * margin + 1 window positions, v1's
// byte shifted against the buffer.  C = (4 + sv_offset % 8) / 8 is the first byte
// visited; out starts zeroed.  margin <= 31 and sv_offset / 8 <= 31 - margin keep
// every shift below 64 (checked by the callers).
template <int C>
__device__ __forceinline__ bool bitmargin_steps(const uint32_t a[8], const uint32_t b[8], uint32_t margin,
                                                uint32_t byte_off, uint32_t bit_off, uint32_t out[8])
{
  uint64_t buffer = 0;
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) buffer += (uint64_t)(k == 0 ? bv_byte(b, 0) & 0xF0u : bv_byte(b, k)) << (32 + 8 * k);
#pragma unroll
  for (uint32_t k = 0; k < 8; ++k) out[k] = 0;
#pragma unroll
  for (uint32_t k = C; k < 32; ++k) {
    const uint32_t i = 4 + byte_off + 8 * (k - C);  // the reference's bit index of this step
    if (i >= 254) break;
    uint32_t byte = 0;
    uint64_t b1 = (uint64_t)bv_byte(a, k) << (32 - margin + bit_off);
    for (uint32_t s = 0; s <= 2 * margin; ++s) {
      byte |= (uint32_t)(uint8_t)((b1 & buffer) >> (32 - margin + s));
      b1 <<= 1;
    }
    out[k >> 2] |= byte << ((k & 3) * 8);
    buffer >>= 8;
    if (i < 224) buffer += (uint64_t)bv_byte(b, (k + 4) & 31) << 56;  // i < 224 implies k + 4 < 32
  }
  return (out[0] | out[1] | out[2] | out[3] | out[4] | out[5] | out[6] | out[7]) != 0u;
}

__device__ __forceinline__ bool bitmargin(const uint32_t a[8], const uint32_t b[8], uint32_t margin,
                                          uint32_t sv_offset, uint32_t out[8])
{
  const uint32_t byte_off = sv_offset % 8, bit_off = sv_offset / 8;
  return (4 + byte_off) >= 8 ? bitmargin_steps<1>(a, b, margin, byte_off, bit_off, out)
                             : bitmargin_steps<0>(a, b, margin, byte_off, bit_off, out);
}

// performLeftShift (helpers/BitVector.h:361-400): bytes k .. k+7 in a 64-bit
// buffer, result byte k = (buffer >> shift) & 0xFF, byte 0 keeps bits 4..7 only
__device__ __forceinline__ void left_shift_sv(uint32_t v[8], uint32_t shift)
{
  uint64_t buffer = (uint64_t)v[0] | ((uint64_t)v[1] << 32);
  uint32_t out[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (uint32_t k = 0; k < 32; ++k) {
    uint32_t nb = (uint32_t)(uint8_t)(buffer >> shift);
    if (k == 0) nb &= 0xF0u;
    out[k >> 2] |= nb << ((k & 3) * 8);
    buffer >>= 8;
    if (k + 8 < 32) buffer += (uint64_t)bv_byte(v, k + 8) << 56;
  }
#pragma unroll
  for (uint32_t k = 0; k < 8; ++k) v[k] = out[k];
}

__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __ballot(p); }

// Runs of consecutive lanes with in set and equal (idx, meaning): returns the run
// length for the first lane of each run and 0 for every other lane.  Neighbouring
// points of an image often share a voxel; one shuffle and one ballot find the runs
// (grouping every distinct pair of the wave takes one ballot round per pair), and
// pairs repeated out of order form several runs.
__device__ __forceinline__ uint32_t wave_run_head(bool in, uint64_t idx, uint32_t meaning)
{
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t inm = wave_ballot(in);
  const uint32_t plo = __shfl_up((uint32_t)idx, 1, 64), phi = __shfl_up((uint32_t)(idx >> 32), 1, 64);
  const uint32_t pm = __shfl_up(meaning, 1, 64);
  const bool same = in && lane > 0 && ((inm >> (lane - 1)) & 1ull) != 0ull && plo == (uint32_t)idx &&
                    phi == (uint32_t)(idx >> 32) && pm == meaning;
  const uint64_t cont = wave_ballot(same);
  if (!in || same) return 0u;
  const uint64_t above = lane == 63 ? 0ull : ~cont & (~0ull << (lane + 1));  // the next run starts there
  const uint32_t end = above != 0ull ? (uint32_t)__builtin_ctzll(above) : 64u;
  return end - lane;
}
__device__ __forceinline__ uint32_t wave_or(uint32_t v)
{
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v)
{
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// deterministic sine/cosine: Cody-Waite reduction by pi/2 (three-part
// constant) and the Cephes single-precision minimax polynomials (S. L. Moshier)
// on [-pi/4, pi/4], every step an explicit fmaf.
__host__ __device__ __forceinline__ void sincos_det(float x, float& s, float& c)
{
  const float two_over_pi = 0.636619772367581343f;
  const float p1 = 1.5703125f;
  const float p2 = 4.837512969970703125e-4f;
  const float p3 = 7.54978995489188216e-8f;
  float k = rintf(x * two_over_pi);
  float r = fmaf(-k, p1, x);
  r = fmaf(-k, p2, r);
  r = fmaf(-k, p3, r);
  float r2 = r * r;
  float ps = fmaf(fmaf(-1.9515295891e-4f, r2, 8.3321608736e-3f), r2, -1.6666654611e-1f);
  float sn = fmaf(ps * r2, r, r);
  float pc = fmaf(fmaf(2.443315711809948e-5f, r2, -1.388731625493765e-3f), r2, 4.166664568298827e-2f);
  float cs = fmaf(pc * r2, r2, fmaf(-0.5f, r2, 1.0f));
  int q = ((int)k) & 3;
  if (q == 0) { s = sn; c = cs; }
  else if (q == 1) { s = cs; c = -sn; }
  else if (q == 2) { s = -sn; c = -cs; }
  else { s = -cs; c = sn; }
}

// DH -> 3x4 affine (robot/dh_robot/KinematicLink.cu:24-79, b = 0)
__host__ __device__ __forceinline__ void dh_matrix(const float* p, float value, float* m)
{
  float th = p[0], dd = p[1];
  if (p[4] != 0.0f) dd = p[1] + value; else th = p[0] + value;
  float st, ct, sa, ca;
  sincos_det(th, st, ct);
  sincos_det(p[3], sa, ca);
  float a = p[2];
  m[0] = ct; m[1] = -(st * ca); m[2] = st * sa;    m[3] = a * ct;
  m[4] = st; m[5] = ct * ca;    m[6] = -(ct * sa); m[7] = a * st;
  m[8] = 0.f; m[9] = sa;        m[10] = ca;        m[11] = dd;
}

// 3x4 affine product, fixed fmaf order (helpers/cuda_matrices.h:773-803 for affine matrices)
__host__ __device__ __forceinline__ void affine_mul(const float* x, const float* y, float* out)
{
  float r[12];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j)
      r[4 * i + j] = fmaf(x[4 * i + 2], y[8 + j], fmaf(x[4 * i + 1], y[4 + j], x[4 * i] * y[j]));
    r[4 * i + 3] = fmaf(x[4 * i + 2], y[11], fmaf(x[4 * i + 1], y[7], fmaf(x[4 * i], y[3], x[4 * i + 3])));
  }
#pragma unroll
  for (int k = 0; k < 12; ++k) out[k] = r[k];
}

}  // namespace gvx

struct gvx_map_s : gvx::MapS {};
