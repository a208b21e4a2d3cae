/*
 * gvx_oracle.c -- serial CPU restatement of GPU-Voxels' data-parallel core.
 *
 * TEST INFRASTRUCTURE ONLY (see gvx_oracle.h).  Built with gcc
 * -ffp-contract=off so that every multiply-add below rounds exactly as
 * written; fused operations are spelled out with fmaf()/fma().
 *
 * Reference paths are relative to packages/gpu_voxels/src/gpu_voxels/.
 */
#include "gvx_oracle.h"

#include <limits.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define BV_BYTES 32u
#define M_OCCUPIED 1u
#define M_COLLISION 2u
#define M_SV_START 4u
#define M_SV_END 254u
#define M_UNCERTAIN 129u
#define M_UNDEFINED 255u
#define PROB_UNKNOWN ((int8_t)-128)
#define PROB_MIN (-127)
#define PROB_MAX 127
#define SENSOR_FREE (-10)
#define SENSOR_OCC 72

static uint64_t nvox(const uint32_t d[3]) { return (uint64_t)d[0] * d[1] * d[2]; }

/* ---------------------------------------------------------------------------
 * Quantization: VoxelMapOperations.h:121-131 computes
 *   static_cast<uint32_t>(floor(p / side)).
 * On the GPU that cast saturates (NaN and negatives -> 0, >= 2^32 -> UINT_MAX);
 * on x86 it is undefined.  The oracle states the GPU rule explicitly.  The
 * division is a true IEEE division.
 * ------------------------------------------------------------------------- */
uint32_t gvo_quantize(float p, float side)
{
  float q = p / side;
  float f = floorf(q);
  if (!(f >= 0.0f)) return 0u;
  if (f >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)f;
}

/* signed variant, VoxelMapOperations.h:134-144 (int32 cast, clamped the same way) */
static int32_t quantize_signed(float p, float side)
{
  float f = floorf(p / side);
  if (f != f) return 0;
  if (f <= -2147483648.0f) return INT_MIN;
  if (f >= 2147483648.0f) return INT_MAX;
  return (int32_t)f;
}

static int64_t signed_index(const uint32_t d[3], const int32_t o[3])
{
  /* VoxelMapOperations.h:58-70 getVoxelIndexSigned; evaluated in 64 bits */
  return (int64_t)o[2] * (int64_t)d[0] * (int64_t)d[1] + (int64_t)o[1] * (int64_t)d[0] + (int64_t)o[0];
}

/* ---------------------------------------------------------------------------
 * Deterministic float sine/cosine.  The reference calls the C library
 * (robot/dh_robot/KinematicLink.cu:60-63); library results differ between
 * host and device in the last ulp, which would flip voxel indices.  Both this
 * oracle and the GPU code therefore use one fixed method: Cody-Waite
 * reduction by pi/2 with a three-part constant, then the single-precision
 * minimax polynomials of the Cephes library (S. L. Moshier, sinf.c/cosf.c)
 * on [-pi/4, pi/4], every step an explicit fmaf.
 * ------------------------------------------------------------------------- */
void gvo_sincosf(float x, float* s, float* c)
{
  const float two_over_pi = 0.636619772367581343f;
  const float p1 = 1.5703125f;              /* pi/2 split: exact high part */
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
  float ss, cc;
  switch (q) {
    case 0: ss = sn; cc = cs; break;
    case 1: ss = cs; cc = -sn; break;
    case 2: ss = -sn; cc = -cs; break;
    default: ss = -cs; cc = sn; break;
  }
  *s = ss;
  *c = cc;
}

/* DH -> 3x4 affine, robot/dh_robot/KinematicLink.cu:24-79 (standard DH, b = 0).
   Layout: row-major rows (a11 a12 a13 a14), (a21 ..), (a31 ..). */
void gvo_dh_matrix(float theta, float d, float a, float alpha, int prismatic, float value, float m[12])
{
  float th = theta, dd = d;
  if (prismatic) dd = d + value; else th = theta + value;
  float st, ct, sa, ca;
  gvo_sincosf(th, &st, &ct);
  gvo_sincosf(alpha, &sa, &ca);
  m[0] = ct;  m[1] = -(st * ca); m[2] = st * sa;    m[3] = a * ct;
  m[4] = st;  m[5] = ct * ca;    m[6] = -(ct * sa); m[7] = a * st;
  m[8] = 0.f; m[9] = sa;         m[10] = ca;        m[11] = dd;
}

/* 3x4 affine product with a fixed fmaf order (helpers/cuda_matrices.h:773-803
   restricted to affine matrices whose last row is 0 0 0 1). */
void gvo_affine_mul(const float x[12], const float y[12], float out[12])
{
  float r[12];
  for (int i = 0; i < 3; ++i) {
    const float* xi = x + 4 * i;
    for (int j = 0; j < 3; ++j)
      r[4 * i + j] = fmaf(xi[2], y[8 + j], fmaf(xi[1], y[4 + j], xi[0] * y[j]));
    r[4 * i + 3] = fmaf(xi[2], y[11], fmaf(xi[1], y[7], fmaf(xi[0], y[3], xi[3])));
  }
  memcpy(out, r, sizeof(r));
}

/* p' = M p, helpers/cuda_matrices.h:741-747, evaluated as
   fmaf(a13, z, fmaf(a12, y, fmaf(a11, x, a14))) per row. */
void gvo_transform_points(const float m[12], const float* pts, uint64_t n, float* out)
{
  for (uint64_t i = 0; i < n; ++i) {
    float x = pts[3 * i], y = pts[3 * i + 1], z = pts[3 * i + 2];
    for (int r = 0; r < 3; ++r)
      out[3 * i + r] = fmaf(m[4 * r + 2], z, fmaf(m[4 * r + 1], y, fmaf(m[4 * r], x, m[4 * r + 3])));
  }
}

/* ProbabilisticVoxel::floatToProbability, voxel/ProbabilisticVoxel.h:39-45 */
int8_t gvo_float_to_probability(float v)
{
  float t = v < 1.0f ? v : 1.0f;
  t = t > 0.0f ? t : 0.0f;
  t = (t * ((float)PROB_MAX - (float)PROB_MIN)) + (float)PROB_MIN;
  return (int8_t)t;
}

/* SVCollider::floatToProbability, voxel/SVCollider.hpp:133-137 (no clamp) */
static int8_t sv_float_to_probability(float v)
{
  float t = (v * ((float)PROB_MAX - (float)PROB_MIN)) + (float)PROB_MIN;
  if (t >= 127.0f) return 127;
  if (t <= -128.0f) return -128;
  return (int8_t)t;
}

/* ProbabilisticVoxel::updateOccupancy, voxel/ProbabilisticVoxel.hpp:43-56 */
int8_t gvo_prob_update(int8_t cur, int8_t delta)
{
  if (cur == PROB_UNKNOWN) return (int8_t)(0 + (int32_t)delta);
  int32_t v = (int32_t)cur + (int32_t)delta;
  if (v < PROB_MIN) v = PROB_MIN;
  if (v > PROB_MAX) v = PROB_MAX;
  return (int8_t)v;
}

/* ProbabilisticVoxel::insert, voxel/ProbabilisticVoxel.hpp:79-101 */
int8_t gvo_prob_insert(int8_t cur, uint32_t meaning)
{
  switch (meaning) {
    case 0: return gvo_prob_update(cur, SENSOR_FREE);
    case M_COLLISION:
    case M_OCCUPIED: return gvo_prob_update(cur, SENSOR_OCC);
    case 3:
    case M_UNDEFINED: return PROB_UNKNOWN;
    default:
      if (meaning <= M_SV_END && meaning >= M_SV_START)
        return gvo_prob_update(cur, (int8_t)((int32_t)meaning - (int32_t)M_UNCERTAIN));
      return PROB_UNKNOWN;
  }
}

/* BitVector::noneButEmpty, helpers/BitVector.h:183-198 */
static int bv_none_but_empty(const uint8_t* b)
{
  if (b[0] & 0xFE) return 0;
  for (unsigned i = 1; i < BV_BYTES; ++i)
    if (b[i]) return 0;
  return 1;
}
static int bv_is_zero(const uint8_t* b)
{
  for (unsigned i = 0; i < BV_BYTES; ++i)
    if (b[i]) return 0;
  return 1;
}
static void bv_set(uint8_t* b, uint32_t i) { b[i >> 3] |= (uint8_t)(1u << (i & 7)); }

/* bitMarginCollisionCheck, helpers/BitVector.h:415-471, restated literally
   (64-bit sliding buffer over v2, SV bits only).  Bytes the loop does not
   visit keep their value in `out`. */
int gvo_bitmargin_collision(const uint8_t v1[32], const uint8_t v2[32], uint8_t out[32],
                            uint32_t margin, uint32_t sv_offset)
{
  const unsigned half = 32;
  uint64_t buffer = 0;
  for (unsigned b = 0; b < 4; ++b) {
    uint8_t in = v2[b];
    if (b == 0) in &= 0xF0;
    buffer += (uint64_t)in << (half + b * 8);
  }
  uint8_t byte_offset = (uint8_t)(sv_offset % 8);
  uint8_t bit_offset = (uint8_t)(sv_offset / 8);
  for (uint32_t i = M_SV_START + byte_offset; i < M_SV_END; i += 8) {
    uint8_t byte = 0;
    uint64_t b1 = (uint64_t)v1[i >> 3] << (half - margin + bit_offset);
    for (uint32_t s = 0; s <= 2 * margin; ++s) {
      byte |= (uint8_t)((b1 & buffer) >> (half - margin + s));
      b1 <<= 1;
    }
    out[i >> 3] = byte;
    buffer >>= 8;
    if (i < 256 - half) buffer += (uint64_t)v2[(i + half) >> 3] << 56;
  }
  return !bv_is_zero(out);
}

/* DistanceVoxel "uninitialised" value: x = y = z = 1023 (helpers/common_defines.h:113,123).
   Extension: grids with an axis of 1024 use 0xFFFFFFFF (z field 4095, never a
   valid coordinate) because 1023 is a valid coordinate there. */
uint32_t gvo_dist_uninit(const uint32_t d[3])
{
  if (d[0] <= 1023 && d[1] <= 1023 && d[2] <= 1023) return 0x3FFFFFFFu;
  return 0xFFFFFFFFu;
}

static uint32_t dist_pack(uint32_t x, uint32_t y, uint32_t z) { return x | (y << 10) | (z << 20); }

/* DistanceVoxel::squaredObstacleDistance, voxel/DistanceVoxel.hpp:38-55 */
int32_t gvo_dist_sq(uint32_t raw, int32_t x, int32_t y, int32_t z, uint32_t uninit)
{
  if (raw == uninit) return INT_MAX;
  int32_t ox = (int32_t)(raw & 0x3FF), oy = (int32_t)((raw >> 10) & 0x3FF), oz = (int32_t)(raw >> 20);
  int32_t dx = x - ox, dy = y - oy, dz = z - oz;
  return dx * dx + dy * dy + dz * dz;
}

/* ---------------------------------------------------------------------------
 * Dense maps
 * ------------------------------------------------------------------------- */

/* clearMap: TemplateVoxelMap.hpp:197-271 (0 / UNKNOWN / uninitialised) */
void gvo_map_clear(int type, void* map, const uint32_t d[3])
{
  uint64_t n = nvox(d);
  if (type == GVO_BITVECTOR_MAP) memset(map, 0, n * BV_BYTES);
  else if (type == GVO_PROB_MAP) memset(map, 0x80, n);
  else {
    uint32_t u = gvo_dist_uninit(d);
    uint32_t* m = (uint32_t*)map;
    for (uint64_t i = 0; i < n; ++i) m[i] = u;
  }
}

static void voxel_insert(int type, void* map, uint64_t idx, uint32_t x, uint32_t y, uint32_t z,
                         uint32_t meaning, uint32_t uninit)
{
  if (type == GVO_BITVECTOR_MAP) {
    bv_set((uint8_t*)map + idx * BV_BYTES, meaning);       /* BitVoxel.hpp:60-65 */
  } else if (type == GVO_PROB_MAP) {
    int8_t* m = (int8_t*)map;
    m[idx] = gvo_prob_insert(m[idx], meaning);
  } else {
    /* DistanceVoxel::insert, DistanceVoxel.hpp:128-139 */
    ((uint32_t*)map)[idx] = (meaning == M_OCCUPIED) ? dist_pack(x, y, z) : uninit;
  }
}

/* kernelInsertGlobalPointCloud, VoxelMapOperations.hpp:241-300.  Serial
   semantics: every point is applied once, in input order. */
void gvo_map_insert_points(int type, void* map, const uint32_t d[3], float side,
                           const float* pts, uint64_t n, uint32_t meaning, int* outside)
{
  uint32_t u = gvo_dist_uninit(d);
  int out = 0;
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t x = gvo_quantize(pts[3 * i], side), y = gvo_quantize(pts[3 * i + 1], side),
             z = gvo_quantize(pts[3 * i + 2], side);
    if (x < d[0] && y < d[1] && z < d[2])
      voxel_insert(type, map, ((uint64_t)z * d[1] + y) * d[0] + x, x, y, z, meaning, u);
    else
      out = 1;
  }
  if (outside) *outside = out;
}

/* gvo_map_insert_points over `threads` OpenMP threads with the same result: the
   points' voxels are found in parallel, then each thread applies, in point order,
   the points that fall in its contiguous range of voxels (updates of one voxel keep
   their order, so saturating Prob updates agree with the serial loop). */
void gvo_map_insert_points_mt(int type, void* map, const uint32_t d[3], float side,
                              const float* pts, uint64_t n, uint32_t meaning, int* outside, int threads)
{
  const uint32_t u = gvo_dist_uninit(d);
  const uint64_t nv = nvox(d);
  uint64_t* vid = (uint64_t*)malloc((n ? n : 1) * sizeof(uint64_t));
  if (!vid) return;
  int out = 0;
  if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(static) reduction(| : out)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const uint32_t x = gvo_quantize(pts[3 * i], side), y = gvo_quantize(pts[3 * i + 1], side),
                   z = gvo_quantize(pts[3 * i + 2], side);
    if (x < d[0] && y < d[1] && z < d[2]) {
      vid[i] = ((uint64_t)z * d[1] + y) * d[0] + x;
    } else {
      vid[i] = UINT64_MAX;
      out = 1;
    }
  }
#pragma omp parallel num_threads(threads)
  {
    const uint64_t t = (uint64_t)omp_get_thread_num(), nt = (uint64_t)omp_get_num_threads();
    const uint64_t lo = nv * t / nt, hi = nv * (t + 1) / nt;
    for (uint64_t i = 0; i < n; ++i) {
      const uint64_t v = vid[i];
      if (v >= lo && v < hi)
        voxel_insert(type, map, v, (uint32_t)(v % d[0]), (uint32_t)((v / d[0]) % d[1]),
                     (uint32_t)(v / ((uint64_t)d[0] * d[1])), meaning, u);
    }
  }
  free(vid);
  if (outside) *outside = out;
}

/* kernelInsertCoordinateTuples, VoxelMapOperations.hpp:302-351 */
void gvo_map_insert_coords(int type, void* map, const uint32_t d[3], const uint32_t* c,
                           uint64_t n, uint32_t meaning, int* outside)
{
  uint32_t u = gvo_dist_uninit(d);
  int out = 0;
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t x = c[3 * i], y = c[3 * i + 1], z = c[3 * i + 2];
    if (x < d[0] && y < d[1] && z < d[2])
      voxel_insert(type, map, ((uint64_t)z * d[1] + y) * d[0] + x, x, y, z, meaning, u);
    else
      out = 1;
  }
  if (outside) *outside = out;
}

/* kernelInsertMetaPointCloud (one meaning / per-cloud meanings),
   VoxelMapOperations.hpp:440-575.  `meanings` may be NULL. */
void gvo_map_insert_meta(int type, void* map, const uint32_t d[3], float side, const float* pts,
                         const uint32_t* sizes, uint32_t ncl, const uint32_t* meanings,
                         uint32_t meaning, int* outside)
{
  uint64_t off = 0;
  int out = 0;
  for (uint32_t c = 0; c < ncl; ++c) {
    int o = 0;
    gvo_map_insert_points(type, map, d, side, pts + 3 * off, sizes[c], meanings ? meanings[c] : meaning, &o);
    out |= o;
    off += sizes[c];
  }
  if (outside) *outside = out;
}

/* BitVoxelMap::insertMetaPointCloudWithSelfCollisionCheck (voxelmap/BitVoxelMap.hpp:228-341,
   kernelInsertMetaPointCloudSelfCollCheck VoxelMapOperations.hpp:582-638), BitVector maps.
   Sub-clouds are applied one after another (the reference launches one kernel per
   sub-cloud).  Within a sub-cloud every point tests the voxel as it was before the
   sub-cloud: masked = voxel & mask[c]; when masked holds a bit other than eBVM_FREE,
   colliding |= masked, colliding gets meaning[c] and the voxel gets eBVM_COLLISION;
   every point in the map then sets meaning[c].  (The reference's points of one
   sub-cloud run concurrently, so a mask that keeps the cloud's own meaning or
   eBVM_COLLISION makes its outcome depend on thread order; with the default masks,
   which clear both, any order gives this result.)  meanings NULL: SV start + c;
   masks NULL: all bits but eBVM_COLLISION and meaning[c].  Returns colliding != 0. */
int gvo_map_insert_meta_selfcoll(uint8_t* map, const uint32_t d[3], float side, const float* pts,
                                 const uint32_t* sizes, uint32_t ncl, const uint32_t* meanings,
                                 const uint8_t* masks, uint8_t colliding[BV_BYTES], int* outside)
{
  uint64_t off = 0;
  int out = 0;
  memset(colliding, 0, BV_BYTES);
  for (uint32_t c = 0; c < ncl; ++c) {
    const uint32_t mc = meanings ? meanings[c] : M_SV_START + c;
    uint8_t mask[BV_BYTES];
    if (masks) memcpy(mask, masks + (uint64_t)c * BV_BYTES, BV_BYTES);
    else {
      memset(mask, 0xFF, BV_BYTES);
      mask[M_COLLISION >> 3] &= (uint8_t)~(1u << (M_COLLISION & 7));
      mask[mc >> 3] &= (uint8_t)~(1u << (mc & 7));
    }
    const uint64_t n = sizes[c];
    uint8_t* hit = (uint8_t*)calloc(n ? n : 1, 1);
    for (uint64_t i = 0; i < n; ++i) {  /* tests against the map before this sub-cloud */
      const float* p = pts + 3 * (off + i);
      uint32_t x = gvo_quantize(p[0], side), y = gvo_quantize(p[1], side), z = gvo_quantize(p[2], side);
      if (!(x < d[0] && y < d[1] && z < d[2])) { out = 1; continue; }
      const uint8_t* v = map + (((uint64_t)z * d[1] + y) * d[0] + x) * BV_BYTES;
      uint8_t masked[BV_BYTES];
      for (unsigned k = 0; k < BV_BYTES; ++k) masked[k] = v[k] & mask[k];
      if (!bv_none_but_empty(masked)) {
        for (unsigned k = 0; k < BV_BYTES; ++k) colliding[k] |= masked[k];
        colliding[mc >> 3] |= (uint8_t)(1u << (mc & 7));
        hit[i] = 1;
      }
    }
    for (uint64_t i = 0; i < n; ++i) {  /* then the sub-cloud's inserts */
      const float* p = pts + 3 * (off + i);
      uint32_t x = gvo_quantize(p[0], side), y = gvo_quantize(p[1], side), z = gvo_quantize(p[2], side);
      if (!(x < d[0] && y < d[1] && z < d[2])) continue;
      uint8_t* v = map + (((uint64_t)z * d[1] + y) * d[0] + x) * BV_BYTES;
      if (hit[i]) v[M_COLLISION >> 3] |= (uint8_t)(1u << (M_COLLISION & 7));
      v[mc >> 3] |= (uint8_t)(1u << (mc & 7));
    }
    free(hit);
    off += n;
  }
  if (outside) *outside = out;
  for (unsigned k = 0; k < BV_BYTES; ++k)
    if (colliding[k]) return 1;
  return 0;
}

/* DefaultCollider, voxel/DefaultCollider.hpp:35-93 */
static int default_collide(int ta, const void* a, uint64_t ia, int tb, const void* b, uint64_t ib, int8_t T)
{
  int occ_a, occ_b;
  if (ta == GVO_BITVECTOR_MAP) occ_a = !bv_none_but_empty((const uint8_t*)a + ia * BV_BYTES);
  else if (ta == GVO_PROB_MAP) occ_a = ((const int8_t*)a)[ia] >= T;
  else return 0;
  if (tb == GVO_BITVECTOR_MAP) occ_b = !bv_none_but_empty((const uint8_t*)b + ib * BV_BYTES);
  else if (tb == GVO_PROB_MAP) occ_b = ((const int8_t*)b)[ib] >= T;
  else return 0;
  return occ_a && occ_b;
}

/* collideWith: collisionCheckWithCounterRelativeTransform (TemplateVoxelMap.hpp:590-621)
   + kernelCollideVoxelMapsDebug (VoxelMapOperations.hpp:129-184).  this[i + off]
   is compared with other[i]; the reference reads outside the buffer when
   i + off leaves [0, N); the oracle skips those i.  Colliding voxels of *this*
   receive insert(eBVM_COLLISION) (bit 2 / +72 log-odds).  The count is the
   true 64-bit count (the reference sums uint16 per-block counters). */
uint64_t gvo_map_collide(int ta, void* a, int tb, const void* b, const uint32_t d[3],
                         float thr, const int32_t offset[3])
{
  uint64_t n = nvox(d), cnt = 0;
  int64_t off = offset ? signed_index(d, offset) : 0;
  int8_t T = gvo_float_to_probability(thr);
  for (uint64_t i = 0; i < n; ++i) {
    int64_t j = (int64_t)i + off;
    if (j < 0 || j >= (int64_t)n) continue;
    if (default_collide(ta, a, (uint64_t)j, tb, b, i, T)) {
      ++cnt;
      if (ta == GVO_BITVECTOR_MAP) bv_set((uint8_t*)a + (uint64_t)j * BV_BYTES, M_COLLISION);
      else if (ta == GVO_PROB_MAP) ((int8_t*)a)[j] = gvo_prob_insert(((int8_t*)a)[j], M_COLLISION);
    }
  }
  return cnt;
}

/* gvo_map_collide over `threads` OpenMP threads: voxels are independent (each
   voxel of *this* is compared with one voxel of other and only it is marked), so
   the count and the marked map equal the serial ones. */
uint64_t gvo_map_collide_mt(int ta, void* a, int tb, const void* b, const uint32_t d[3],
                            float thr, const int32_t offset[3], int threads)
{
  const uint64_t n = nvox(d);
  const int64_t off = offset ? signed_index(d, offset) : 0;
  const int8_t T = gvo_float_to_probability(thr);
  uint64_t cnt = 0;
  if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(static) reduction(+ : cnt)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const int64_t j = i + off;
    if (j < 0 || j >= (int64_t)n) continue;
    if (default_collide(ta, a, (uint64_t)j, tb, b, (uint64_t)i, T)) {
      ++cnt;
      if (ta == GVO_BITVECTOR_MAP) bv_set((uint8_t*)a + (uint64_t)j * BV_BYTES, M_COLLISION);
      else if (ta == GVO_PROB_MAP) ((int8_t*)a)[j] = gvo_prob_insert(((int8_t*)a)[j], M_COLLISION);
    }
  }
  return cnt;
}

/* collideWithTypes: BitVoxelMap.hpp:92-224 + kernelCollideVoxelMapsBitvector
   (VoxelMapOperations.hpp:187-239) with SVCollider (voxel/SVCollider.hpp).  The
   reference keeps only the last colliding mask per thread (cache[i] = temp);
   the oracle ORs the masks of all colliding voxels. */
uint64_t gvo_map_collide_types(int ta, void* a, int tb, const void* b, const uint32_t d[3],
                               float thr, uint8_t mask[32])
{
  return gvo_map_collide_sv(ta, a, tb, b, d, thr, 0, 0, mask);
}

/* collisionCheckBitvector with SVCollider(thr, window) (BitVoxelMap.hpp:90-195,
   kernelCollideVoxelMapsBitvector VoxelMapOperations.hpp:186-239,
   SVCollider.hpp:40-110): Bit x Bit voxels collide by bitMarginCollisionCheck
   with margin = window and the given sv_offset (a zeroed result per voxel);
   Bit x Prob as collideWithTypes.  Mask = OR over every colliding voxel. */
uint64_t gvo_map_collide_sv(int ta, void* a, int tb, const void* b, const uint32_t d[3],
                            float thr, uint32_t window, uint32_t sv_offset, uint8_t mask[32])
{
  uint64_t n = nvox(d), cnt = 0;
  int8_t T = sv_float_to_probability(thr);
  uint8_t acc[32];
  memset(acc, 0, 32);
  if (ta != GVO_BITVECTOR_MAP) return 0;
  uint8_t* A = (uint8_t*)a;
  for (uint64_t i = 0; i < n; ++i) {
    uint8_t* va = A + i * BV_BYTES;
    int hit = 0;
    if (tb == GVO_BITVECTOR_MAP) {
      uint8_t tmp[32];
      memset(tmp, 0, 32);
      hit = gvo_bitmargin_collision(va, (const uint8_t*)b + i * BV_BYTES, tmp, window, sv_offset);
      if (hit)
        for (int k = 0; k < 32; ++k) acc[k] |= tmp[k];
    } else if (tb == GVO_PROB_MAP) {
      hit = (((const int8_t*)b)[i] >= T) && !bv_none_but_empty(va);
      if (hit)
        for (int k = 0; k < 32; ++k) acc[k] |= va[k];
    }
    if (hit) {
      ++cnt;
      bv_set(va, M_COLLISION);
    }
  }
  if (mask) memcpy(mask, acc, 32);
  return cnt;
}

/* kernelClearVoxelMap(bits), VoxelMapOperations.hpp:59-72 */
void gvo_map_clear_bits(void* map, const uint32_t d[3], const uint8_t bits[32])
{
  uint64_t n = nvox(d);
  uint8_t* m = (uint8_t*)map;
  for (uint64_t i = 0; i < n; ++i)
    for (int k = 0; k < 32; ++k) m[i * BV_BYTES + k] &= (uint8_t)~bits[k];
}

/* ---------------------------------------------------------------------------
 * Voxel lists (TemplateVoxelList.hpp / BitVoxelList.hpp, BitVectorVoxel payload)
 * ------------------------------------------------------------------------- */

typedef struct { uint32_t id; uint64_t pos; } keypos_t;
static int cmp_keypos(const void* a, const void* b)
{
  const keypos_t* x = (const keypos_t*)a; const keypos_t* y = (const keypos_t*)b;
  if (x->id != y->id) return x->id < y->id ? -1 : 1;
  return x->pos < y->pos ? -1 : (x->pos > y->pos ? 1 : 0);
}

/* make_unique, TemplateVoxelList.hpp:143-212: stable sort by id, merge each run
   of equal ids into its first element with BitVoxel::reduce (OR), drop the rest. */
void gvo_list_make_unique(uint32_t* ids, uint32_t* coords, uint8_t* bits, uint64_t* len)
{
  uint64_t n = *len;
  if (n == 0) return;
  keypos_t* kp = (keypos_t*)malloc(n * sizeof(keypos_t));
  uint32_t* oid = (uint32_t*)malloc(n * 4);
  uint32_t* oc = (uint32_t*)malloc(n * 12);
  uint8_t* ob = (uint8_t*)malloc(n * BV_BYTES);
  for (uint64_t i = 0; i < n; ++i) { kp[i].id = ids[i]; kp[i].pos = i; }
  qsort(kp, n, sizeof(keypos_t), cmp_keypos);
  uint64_t m = 0;
  for (uint64_t i = 0; i < n;) {
    uint64_t j = i;
    uint64_t p = kp[i].pos;
    oid[m] = ids[p];
    memcpy(oc + 3 * m, coords + 3 * p, 12);
    memcpy(ob + m * BV_BYTES, bits + p * BV_BYTES, BV_BYTES);
    for (j = i + 1; j < n && kp[j].id == kp[i].id; ++j)
      for (int k = 0; k < 32; ++k) ob[m * BV_BYTES + k] |= bits[kp[j].pos * BV_BYTES + k];
    ++m;
    i = j;
  }
  memcpy(ids, oid, m * 4);
  memcpy(coords, oc, m * 12);
  memcpy(bits, ob, m * BV_BYTES);
  *len = m;
  free(kp); free(oid); free(oc); free(ob);
}

/* remove_out_of_bounds, TemplateVoxelList.hpp:297-341 (stable) */
static void list_remove_oob(uint32_t* ids, uint32_t* coords, uint8_t* bits, uint64_t start,
                            uint64_t* len, const uint32_t d[3])
{
  uint64_t w = start;
  for (uint64_t i = start; i < *len; ++i) {
    if (coords[3 * i] >= d[0] || coords[3 * i + 1] >= d[1] || coords[3 * i + 2] >= d[2]) continue;
    if (w != i) {
      ids[w] = ids[i];
      memcpy(coords + 3 * w, coords + 3 * i, 12);
      memcpy(bits + w * BV_BYTES, bits + i * BV_BYTES, BV_BYTES);
    }
    ++w;
  }
  *len = w;
}

static uint32_t id_of(const uint32_t d[3], uint32_t x, uint32_t y, uint32_t z)
{
  /* getVoxelIndexUnsigned in uint32 arithmetic (wraps like the reference) */
  return z * d[0] * d[1] + y * d[0] + x;
}

/* insertPointCloud / insertMetaPointCloud of a list: TemplateVoxelList.hpp:363-514
   with kernels VoxelListOperations.hpp:41-140; then remove_out_of_bounds and
   make_unique.  cloud_sizes/meanings may be NULL (one cloud, one meaning). */
void gvo_list_insert_points(uint32_t* ids, uint32_t* coords, uint8_t* bits, uint64_t* len,
                            const uint32_t d[3], float side, const float* pts, uint64_t n,
                            const uint32_t* sizes, uint32_t ncl, const uint32_t* meanings,
                            uint32_t meaning)
{
  uint64_t base = *len;
  uint32_t cl = 0;
  uint64_t bound = sizes ? sizes[0] : n;
  for (uint64_t i = 0; i < n; ++i) {
    if (sizes) {
      while (i >= bound && cl + 1 < ncl) { ++cl; bound += sizes[cl]; }
    }
    uint32_t x = gvo_quantize(pts[3 * i], side), y = gvo_quantize(pts[3 * i + 1], side),
             z = gvo_quantize(pts[3 * i + 2], side);
    uint64_t k = base + i;
    coords[3 * k] = x; coords[3 * k + 1] = y; coords[3 * k + 2] = z;
    ids[k] = id_of(d, x, y, z);
    memset(bits + k * BV_BYTES, 0, BV_BYTES);
    bv_set(bits + k * BV_BYTES, (sizes && meanings) ? meanings[cl] : meaning);
  }
  *len = base + n;
  list_remove_oob(ids, coords, bits, base, len, d);
  gvo_list_make_unique(ids, coords, bits, len);
}

/* insertCoordinateList, TemplateVoxelList.hpp:395-440 */
void gvo_list_insert_coords(uint32_t* ids, uint32_t* coords, uint8_t* bits, uint64_t* len,
                            const uint32_t d[3], const uint32_t* in, uint64_t n, uint32_t meaning)
{
  uint64_t base = *len;
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t k = base + i;
    memcpy(coords + 3 * k, in + 3 * i, 12);
    ids[k] = id_of(d, in[3 * i], in[3 * i + 1], in[3 * i + 2]);
    memset(bits + k * BV_BYTES, 0, BV_BYTES);
    bv_set(bits + k * BV_BYTES, meaning);
  }
  *len = base + n;
  list_remove_oob(ids, coords, bits, base, len, d);
  gvo_list_make_unique(ids, coords, bits, len);
}

/* merge(list), TemplateVoxelList.hpp:878-916: append, shift by offset
   (coord += offset, id += signed index), optionally overwrite the meaning,
   make_unique.  No bounds filtering after the shift (as the reference). */
void gvo_list_append_list(uint32_t* ids, uint32_t* coords, uint8_t* bits, uint64_t* len,
                          const uint32_t d[3], const uint32_t* oids, const uint32_t* ocoords,
                          const uint8_t* obits, uint64_t on, const int32_t off[3],
                          int has_meaning, uint32_t meaning)
{
  uint64_t base = *len;
  int64_t ao = off ? signed_index(d, off) : 0;
  for (uint64_t i = 0; i < on; ++i) {
    uint64_t k = base + i;
    ids[k] = (uint32_t)((int64_t)oids[i] + ao);
    for (int c = 0; c < 3; ++c) coords[3 * k + c] = ocoords[3 * i + c] + (uint32_t)(off ? off[c] : 0);
    if (has_meaning) { memset(bits + k * BV_BYTES, 0, BV_BYTES); bv_set(bits + k * BV_BYTES, meaning); }
    else memcpy(bits + k * BV_BYTES, obits + i * BV_BYTES, BV_BYTES);
  }
  *len = base + on;
  gvo_list_make_unique(ids, coords, bits, len);
}

/* merge(BitVectorVoxelMap), TemplateVoxelList.hpp:969-1030: occupied voxels
   (!noneButEmpty) in ascending index order, their bits, optional offset and
   meaning, make_unique.  Returns the number of voxels appended. */
uint64_t gvo_list_append_map(uint32_t* ids, uint32_t* coords, uint8_t* bits, uint64_t* len,
                             const uint32_t d[3], const uint8_t* bitmap, const int32_t off[3],
                             int has_meaning, uint32_t meaning)
{
  uint64_t n = nvox(d), base = *len, k = base;
  int64_t ao = off ? signed_index(d, off) : 0;
  for (uint64_t i = 0; i < n; ++i) {
    const uint8_t* v = bitmap + i * BV_BYTES;
    if (bv_none_but_empty(v)) continue;
    uint32_t x = (uint32_t)(i % d[0]), y = (uint32_t)((i / d[0]) % d[1]), z = (uint32_t)(i / ((uint64_t)d[0] * d[1]));
    ids[k] = (uint32_t)((int64_t)i + ao);
    coords[3 * k] = x + (uint32_t)(off ? off[0] : 0);
    coords[3 * k + 1] = y + (uint32_t)(off ? off[1] : 0);
    coords[3 * k + 2] = z + (uint32_t)(off ? off[2] : 0);
    if (has_meaning) { memset(bits + k * BV_BYTES, 0, BV_BYTES); bv_set(bits + k * BV_BYTES, meaning); }
    else memcpy(bits + k * BV_BYTES, v, BV_BYTES);
    ++k;
  }
  *len = k;
  gvo_list_make_unique(ids, coords, bits, len);
  return k - base;
}

/* list x map collision, VoxelListOperations.hpp:146-334 via
   TemplateVoxelList.hpp:1170-1213 and BitVoxelList.hpp:63-258.
   mode 0 collideWith (DefaultCollider), 1 collideWithTypes (isOccupied(thr),
   OR of list bits), 2 collideWithTypeMask.  The map voxel is map[id + off]
   (signed linear offset, in range [0, N) or skipped).  Colliding entries get
   bit eBVM_COLLISION.  The mask is taken before that bit is set. */
uint64_t gvo_list_collide_map(uint32_t* ids, uint8_t* bits, uint64_t len, int mt, const void* map,
                              const uint32_t d[3], float thr, const int32_t off[3], int mode,
                              const uint8_t typemask[32], uint8_t mask_out[32])
{
  uint64_t n = nvox(d), cnt = 0;
  int64_t ao = off ? signed_index(d, off) : 0;
  int8_t T = gvo_float_to_probability(thr);
  uint8_t acc[32];
  memset(acc, 0, 32);
  for (uint64_t i = 0; i < len; ++i) {
    int64_t j = (int64_t)ids[i] + ao;
    if (j < 0 || j >= (int64_t)n) continue;
    uint8_t* e = bits + i * BV_BYTES;
    int map_occ;
    if (mt == GVO_BITVECTOR_MAP) map_occ = !bv_none_but_empty((const uint8_t*)map + (uint64_t)j * BV_BYTES);
    else if (mode == 0) map_occ = ((const int8_t*)map)[j] >= T;
    else map_occ = (float)((const int8_t*)map)[j] >= thr;     /* ProbabilisticVoxel.hpp:102-106 */
    int hit;
    if (mode == 0) hit = map_occ && !bv_none_but_empty(e);
    else if (mode == 1) hit = map_occ;
    else {
      hit = 0;
      if (map_occ)
        for (int k = 0; k < 32; ++k) if (typemask[k] & e[k]) { hit = 1; break; }
    }
    if (hit) {
      ++cnt;
      if (mode == 1) for (int k = 0; k < 32; ++k) acc[k] |= e[k];
      bv_set(e, M_COLLISION);
    }
  }
  if (mask_out) memcpy(mask_out, acc, 32);
  return cnt;
}

static int bsearch_u32(const uint32_t* a, uint64_t n, uint32_t key, uint64_t* pos)
{
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    uint64_t mid = (lo + hi) / 2;
    if (a[mid] < key) lo = mid + 1; else hi = mid;
  }
  *pos = lo;
  return lo < n && a[lo] == key;
}

/* list x list: collideVoxellists (TemplateVoxelList.hpp:231-289) counts ids of A
   found in B; collideWithTypes(list) (BitVoxelList.hpp:94-118) ORs both lists'
   bits over the matches.  Both lists sorted and unique. */
uint64_t gvo_list_collide_list(const uint32_t* ia, const uint8_t* ba, uint64_t na,
                               const uint32_t* ib, const uint8_t* bb, uint64_t nb, uint8_t mask[32])
{
  uint64_t cnt = 0, p;
  uint8_t acc[32];
  memset(acc, 0, 32);
  for (uint64_t i = 0; i < na; ++i) {
    if (bsearch_u32(ib, nb, ia[i], &p)) {
      ++cnt;
      for (int k = 0; k < 32; ++k) acc[k] |= (uint8_t)(ba[i * BV_BYTES + k] | bb[p * BV_BYTES + k]);
    }
  }
  if (mask) memcpy(mask, acc, 32);
  return cnt;
}

/* ---------------------------------------------------------------------------
 * Distance maps
 * ------------------------------------------------------------------------- */

/* mergeOccupied, DistanceVoxelMap.hpp:109-136 with DistanceVoxelMap.h:92-134:
   voxel i receives coords(i)+offset where prob[i] > floatToProbability(thr). */
void gvo_dist_merge_occupied(uint32_t* dist, const int8_t* prob, const uint32_t d[3],
                             const uint32_t off[3], float thr)
{
  uint64_t n = nvox(d);
  int8_t T = gvo_float_to_probability(thr);
  for (uint64_t i = 0; i < n; ++i) {
    if (prob[i] > T) {
      uint32_t x = (uint32_t)(i % d[0]), y = (uint32_t)((i / d[0]) % d[1]), z = (uint32_t)(i / ((uint64_t)d[0] * d[1]));
      dist[i] = dist_pack(x + off[0], y + off[1], z + off[2]);
    }
  }
}

/* One PBA proximate-point pass over a line (VoxelMapOperationsPBA.hpp:289-379
   phase 2 and :604-703 phase 3) with the float bisector test of
   devicePBAintersection3DY (:265-276) replaced by the exact integer predicate.
   Sites (pos[k], cost[k]) come in increasing pos; for site a < b the bisector
   meets the line at Y = ((pb^2 - pa^2) + (cb - ca)) / (2 (pb - pa)).
   Phase 2 keeps the stack top t iff Y(s,t) < Y(t,c); phase 3 walks positions
   from L-1 down and pops while y <= Y(second, top).  Writes the chosen site
   index per position (or -1). */
static void pba_line(const int32_t* pos, const int64_t* cost, int nsite, int L,
                     int* stack, int32_t* choice)
{
  int n = 0;
  for (int k = 0; k < nsite; ++k) {
    while (n >= 2) {
      int s = stack[n - 2], t = stack[n - 1];
      int64_t N1 = ((int64_t)pos[t] * pos[t] - (int64_t)pos[s] * pos[s]) + (cost[t] - cost[s]);
      int64_t D1 = 2 * (int64_t)(pos[t] - pos[s]);
      int64_t N2 = ((int64_t)pos[k] * pos[k] - (int64_t)pos[t] * pos[t]) + (cost[k] - cost[t]);
      int64_t D2 = 2 * (int64_t)(pos[k] - pos[t]);
      if (N1 * D2 < N2 * D1) break;
      --n;
    }
    stack[n++] = k;
  }
  if (n == 0) {
    for (int y = 0; y < L; ++y) choice[y] = -1;
    return;
  }
  int top = n - 1;
  for (int y = L - 1; y >= 0; --y) {
    while (top >= 1) {
      int s = stack[top - 1], t = stack[top];
      int64_t N = ((int64_t)pos[t] * pos[t] - (int64_t)pos[s] * pos[s]) + (cost[t] - cost[s]);
      int64_t D = 2 * (int64_t)(pos[t] - pos[s]);
      if ((int64_t)y * D <= N) --top; else break;
    }
    choice[y] = stack[top];
  }
}

/* parallelBanding3D with m1 = m2 = m3 = 1 (DistanceVoxelMap.hpp:284-645),
   restated serially with exact integer predicates.  Every voxel that is not
   "uninitialised" is an obstacle located at its own coordinate.  Phase 1
   floods along z (ties to the lower z, VoxelMapOperationsPBA.hpp:63-124),
   then the proximate-point pass runs along y and finally along x (the
   reference transposes x/y for the last pass, :825-888). */
static void pba_impl(uint32_t* map, const uint32_t d[3], int threads)
{
  const uint32_t X = d[0], Y = d[1], Z = d[2];
  const uint64_t n = nvox(d);
  const uint32_t u = gvo_dist_uninit(d);
  int16_t* zs = (int16_t*)malloc(n * sizeof(int16_t));
  int16_t* ys = (int16_t*)malloc(n * sizeof(int16_t));
  int16_t* zs2 = (int16_t*)malloc(n * sizeof(int16_t));
  uint32_t L = X > Y ? X : Y;
  L = L > Z ? L : Z;
  const uint64_t row = X, layer = (uint64_t)X * Y;
  if (threads < 1) threads = 1;

#pragma omp parallel num_threads(threads)
  {
    /* per-thread line scratch */
    int32_t* pos = (int32_t*)malloc(L * sizeof(int32_t));
    int64_t* cost = (int64_t*)malloc(L * sizeof(int64_t));
    int* stack = (int*)malloc(L * sizeof(int));
    int32_t* choice = (int32_t*)malloc(L * sizeof(int32_t));
    int32_t* site_of = (int32_t*)malloc(L * sizeof(int32_t));

    /* phase 1: z */
#pragma omp for schedule(static)
    for (int64_t yy = 0; yy < (int64_t)Y; ++yy) {
      const uint32_t y = (uint32_t)yy;
      for (uint32_t x = 0; x < X; ++x) {
        int32_t last = -1;
        for (uint32_t z = 0; z < Z; ++z) {
          uint64_t id = z * layer + y * row + x;
          if (map[id] != u) last = (int32_t)z;
          zs[id] = (int16_t)last;
        }
        int32_t next = -1;
        for (int64_t z = (int64_t)Z - 1; z >= 0; --z) {
          uint64_t id = (uint64_t)z * layer + y * row + x;
          if (map[id] != u) next = (int32_t)z;
          int32_t lo = zs[id];
          int32_t pick;
          if (lo < 0) pick = next;
          else if (next < 0) pick = lo;
          else pick = ((int32_t)z - lo <= next - (int32_t)z) ? lo : next;
          zs[id] = (int16_t)pick;
        }
      }
    }

    /* y pass: lines (x, z) */
#pragma omp for schedule(static)
    for (int64_t zz = 0; zz < (int64_t)Z; ++zz) {
      const uint64_t z = (uint64_t)zz;
      for (uint32_t x = 0; x < X; ++x) {
        int ns = 0;
        for (uint32_t y = 0; y < Y; ++y) {
          int16_t s = zs[z * layer + y * row + x];
          if (s >= 0) {
            int64_t dz = (int64_t)s - (int64_t)z;
            pos[ns] = (int32_t)y; cost[ns] = dz * dz; site_of[ns] = (int32_t)y; ++ns;
          }
        }
        pba_line(pos, cost, ns, (int)Y, stack, choice);
        for (uint32_t y = 0; y < Y; ++y) {
          uint64_t id = z * layer + y * row + x;
          if (choice[y] < 0) { ys[id] = -1; zs2[id] = -1; }
          else {
            int32_t sy = site_of[choice[y]];
            ys[id] = (int16_t)sy;
            zs2[id] = zs[z * layer + (uint64_t)sy * row + x];
          }
        }
      }
    }

    /* x pass: lines (y, z) */
#pragma omp for schedule(static)
    for (int64_t zz = 0; zz < (int64_t)Z; ++zz) {
      const uint64_t z = (uint64_t)zz;
      for (uint32_t y = 0; y < Y; ++y) {
        int ns = 0;
        uint64_t base = z * layer + y * row;
        for (uint32_t x = 0; x < X; ++x) {
          int16_t sy = ys[base + x];
          if (sy >= 0) {
            int64_t dy = (int64_t)sy - y, dz = (int64_t)zs2[base + x] - (int64_t)z;
            pos[ns] = (int32_t)x; cost[ns] = dy * dy + dz * dz; site_of[ns] = (int32_t)x; ++ns;
          }
        }
        pba_line(pos, cost, ns, (int)X, stack, choice);
        for (uint32_t x = 0; x < X; ++x) {
          if (choice[x] < 0) map[base + x] = u;
          else {
            uint32_t sx = (uint32_t)site_of[choice[x]];
            map[base + x] = dist_pack(sx, (uint32_t)ys[base + sx], (uint32_t)zs2[base + sx]);
          }
        }
      }
    }
    free(pos); free(cost); free(stack); free(choice); free(site_of);
  }
  free(zs); free(ys); free(zs2);
}

void gvo_pba(uint32_t* map, const uint32_t d[3]) { pba_impl(map, d, 1); }

/* The same passes with the lines of each pass spread over `threads` OpenMP
   threads (lines are independent within a pass; the result is identical). */
void gvo_pba_mt(uint32_t* map, const uint32_t d[3], int threads) { pba_impl(map, d, threads); }

/* Brute-force exact EDT: for every voxel the nearest obstacle voxel, ties
   broken to the lexicographically smallest (x, y, z).  O(N * M); small grids. */
void gvo_edt_bruteforce(const uint32_t* map, const uint32_t d[3], uint32_t* out)
{
  const uint64_t n = nvox(d);
  const uint32_t u = gvo_dist_uninit(d);
  uint64_t m = 0;
  int32_t* ob = (int32_t*)malloc(n * 3 * sizeof(int32_t));
  for (uint32_t x = 0; x < d[0]; ++x)
    for (uint32_t y = 0; y < d[1]; ++y)
      for (uint32_t z = 0; z < d[2]; ++z) {
        uint64_t id = ((uint64_t)z * d[1] + y) * d[0] + x;
        if (map[id] != u) { ob[3 * m] = (int32_t)x; ob[3 * m + 1] = (int32_t)y; ob[3 * m + 2] = (int32_t)z; ++m; }
      }
  for (uint64_t id = 0; id < n; ++id) {
    int32_t x = (int32_t)(id % d[0]), y = (int32_t)((id / d[0]) % d[1]), z = (int32_t)(id / ((uint64_t)d[0] * d[1]));
    int64_t best = INT64_MAX;
    uint64_t bk = 0;
    for (uint64_t k = 0; k < m; ++k) {
      int64_t dx = x - ob[3 * k], dy = y - ob[3 * k + 1], dz = z - ob[3 * k + 2];
      int64_t dd = dx * dx + dy * dy + dz * dz;
      if (dd < best) { best = dd; bk = k; }
    }
    out[id] = (m == 0) ? u : dist_pack((uint32_t)ob[3 * bk], (uint32_t)ob[3 * bk + 1], (uint32_t)ob[3 * bk + 2]);
  }
  free(ob);
}

/* exactDistances3D, DistanceVoxelMap.hpp:208-253 + kernelExactDistances3D
   (VoxelMapOperations.hpp:896-990): per voxel, scan the obstacle point list
   in order (quantized with mapToVoxelsSigned), skipping obstacles at the
   voxel itself and those with a coordinate of 1023; strict < keeps the first
   minimum; the voxel is updated only when that beats its current distance. */
void gvo_exact_distances(uint32_t* map, const uint32_t d[3], float side, const float* obst, uint64_t no)
{
  const uint64_t n = nvox(d);
  const uint32_t u = gvo_dist_uninit(d);
  int32_t* q = (int32_t*)malloc((no ? no : 1) * 3 * sizeof(int32_t));
  for (uint64_t k = 0; k < no; ++k)
    for (int c = 0; c < 3; ++c) q[3 * k + c] = quantize_signed(obst[3 * k + c], side);
  for (uint64_t id = 0; id < n; ++id) {
    int32_t x = (int32_t)(id % d[0]), y = (int32_t)((id / d[0]) % d[1]), z = (int32_t)(id / ((uint64_t)d[0] * d[1]));
    int32_t cur = gvo_dist_sq(map[id], x, y, z, u);
    int32_t best = cur;
    int found = 0;
    int32_t bo[3] = {0, 0, 0};
    if (cur == 0) continue;
    for (uint64_t k = 0; k < no; ++k) {
      const int32_t* o = q + 3 * k;
      if (o[0] == x && o[1] == y && o[2] == z) continue;
      if (o[0] == 1023 || o[1] == 1023 || o[2] == 1023) continue;
      int32_t dx = x - o[0], dy = y - o[1], dz = z - o[2];
      int32_t dd = dx * dx + dy * dy + dz * dz;
      if (dd < best) { best = dd; bo[0] = o[0]; bo[1] = o[1]; bo[2] = o[2]; found = 1; }
    }
    if (found && best < cur) map[id] = (uint32_t)bo[0] | ((uint32_t)bo[1] << 10) | ((uint32_t)bo[2] << 20);
  }
  free(q);
}

/* getSquaredDistances, DistanceVoxelMap.hpp:716-752 */
void gvo_dist_squared(const uint32_t* map, const uint32_t d[3], const uint32_t* idx, uint64_t n, int32_t* out)
{
  const uint32_t u = gvo_dist_uninit(d);
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t id = idx[i];
    int32_t x = (int32_t)(id % d[0]), y = (int32_t)((id / d[0]) % d[1]), z = (int32_t)(id / ((uint64_t)d[0] * d[1]));
    out[i] = gvo_dist_sq(map[id], x, y, z, u);
  }
}

/* ---------------------------------------------------------------------------
 * Robot chain and batched collision
 * ------------------------------------------------------------------------- */

/* KinematicChain::setConfiguration, robot/dh_robot/KinematicChain.cu:93-126:
   cloud i is transformed by base * A_0 * ... * A_{i-1}. */
void gvo_fk_chain(const float* dh, uint32_t nl, const float base[12], const float* q, float* out)
{
  float t[12], a[12];
  memcpy(t, base, sizeof(t));
  for (uint32_t l = 0; l < nl; ++l) {
    memcpy(out + 12 * l, t, sizeof(t));
    const float* p = dh + 5 * l;
    gvo_dh_matrix(p[0], p[1], p[2], p[3], p[4] != 0.0f, q[l], a);
    gvo_affine_mul(t, a, t);
  }
}

static int cmp_u32(const void* a, const void* b)
{
  uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

/* One pose of the reference planner pattern (gvl_ompl_planning/
   gvl_ompl_planner_helper.cpp:169-200): clearMap, FK, insertMetaPointCloud with
   link l -> meaning eBVM_SWEPT_VOLUME_START + l, collideWith(env).  The count
   is the number of distinct robot voxels whose environment voxel is occupied;
   the mask ORs the robot voxel bits over those voxels (the list form of
   collideWithTypes, BitVoxelList.hpp:120-162). */
/* one pose: distinct colliding voxels and the OR of the colliding links' meanings;
   hits (total link points) and T (12 x nl) are the caller's scratch */
static void batch_one_pose(const uint8_t* env, const uint32_t d[3], float side, const float* pts,
                           const uint32_t* sizes, uint32_t nl, const float* dh, const float base[12],
                           const float* qp, uint32_t* hits, float* T, uint32_t* count, uint8_t* mask)
{
  gvo_fk_chain(dh, nl, base, qp, T);
  memset(mask, 0, 32);
  uint64_t nh = 0, off = 0;
  for (uint32_t l = 0; l < nl; ++l) {
    const float* m = T + 12 * l;
    for (uint32_t i = 0; i < sizes[l]; ++i) {
      const float* s = pts + 3 * (off + i);
      float w[3];
      for (int r = 0; r < 3; ++r)
        w[r] = fmaf(m[4 * r + 2], s[2], fmaf(m[4 * r + 1], s[1], fmaf(m[4 * r], s[0], m[4 * r + 3])));
      uint32_t x = gvo_quantize(w[0], side), y = gvo_quantize(w[1], side), z = gvo_quantize(w[2], side);
      if (x >= d[0] || y >= d[1] || z >= d[2]) continue;
      uint64_t v = ((uint64_t)z * d[1] + y) * d[0] + x;
      if (env[v >> 3] & (1u << (v & 7))) {
        hits[nh++] = (uint32_t)v;
        bv_set(mask, M_SV_START + l);
      }
    }
    off += sizes[l];
  }
  qsort(hits, nh, sizeof(uint32_t), cmp_u32);
  uint32_t c = 0;
  for (uint64_t i = 0; i < nh; ++i)
    if (i == 0 || hits[i] != hits[i - 1]) ++c;
  *count = c;
}

void gvo_batch_collide(const uint8_t* env, const uint32_t d[3], float side, const float* pts,
                       const uint32_t* sizes, uint32_t nl, const float* dh, const float base[12],
                       const float* q, uint32_t npose, uint32_t* counts, uint8_t* masks)
{
  uint64_t total = 0;
  for (uint32_t l = 0; l < nl; ++l) total += sizes[l];
  uint32_t* hits = (uint32_t*)malloc((total ? total : 1) * sizeof(uint32_t));
  float* T = (float*)malloc(12 * nl * sizeof(float));
  for (uint32_t p = 0; p < npose; ++p)
    batch_one_pose(env, d, side, pts, sizes, nl, dh, base, q + (uint64_t)p * nl, hits, T, counts + p,
                   masks + 32 * (uint64_t)p);
  free(hits);
  free(T);
}

/* the same over `threads` OpenMP threads (poses are independent; each thread has its
   own scratch), the all-cores CPU baseline */
void gvo_batch_collide_mt(const uint8_t* env, const uint32_t d[3], float side, const float* pts,
                          const uint32_t* sizes, uint32_t nl, const float* dh, const float base[12],
                          const float* q, uint32_t npose, uint32_t* counts, uint8_t* masks, int threads)
{
  uint64_t total = 0;
  for (uint32_t l = 0; l < nl; ++l) total += sizes[l];
#pragma omp parallel num_threads(threads > 0 ? threads : 1)
  {
    uint32_t* hits = (uint32_t*)malloc((total ? total : 1) * sizeof(uint32_t));
    float* T = (float*)malloc(12 * nl * sizeof(float));
#pragma omp for schedule(dynamic, 4)
    for (int64_t p = 0; p < (int64_t)npose; ++p)
      batch_one_pose(env, d, side, pts, sizes, nl, dh, base, q + (uint64_t)p * nl, hits, T, counts + p,
                     masks + 32 * (uint64_t)p);
    free(hits);
    free(T);
  }
}

/* ---------------------------------------------------------------------------
 * Differentiable clearance cost (FP64 restatement; no reference counterpart)
 * d(x): trilinear interpolation of a float grid whose samples sit at voxel
 * centres ((i + 1/2) * side); u = x/side - 1/2 clamped to [0, dim-1].
 * ------------------------------------------------------------------------- */
double gvo_trilinear_d(const float* g, const uint32_t d[3], double side, const double x[3], double grad[3])
{
  int i0[3];
  double t[3], dsc[3];
  for (int a = 0; a < 3; ++a) {
    double u = x[a] / side - 0.5;
    dsc[a] = 1.0 / side;
    double hi = (double)(d[a] - 1);
    if (u < 0.0) { u = 0.0; dsc[a] = 0.0; }
    if (u > hi) { u = hi; dsc[a] = 0.0; }
    int i = (int)floor(u);
    if (i > (int)d[a] - 2) i = (int)d[a] - 2;
    if (i < 0) i = 0;
    i0[a] = i;
    t[a] = u - i;
  }
  double c[2][2][2];
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx)
        c[dz][dy][dx] = g[((uint64_t)(i0[2] + dz) * d[1] + (uint64_t)(i0[1] + dy)) * d[0] + (uint64_t)(i0[0] + dx)];
  double tx = t[0], ty = t[1], tz = t[2];
  double c00 = c[0][0][0] * (1 - tx) + c[0][0][1] * tx, c01 = c[0][1][0] * (1 - tx) + c[0][1][1] * tx;
  double c10 = c[1][0][0] * (1 - tx) + c[1][0][1] * tx, c11 = c[1][1][0] * (1 - tx) + c[1][1][1] * tx;
  double c0 = c00 * (1 - ty) + c01 * ty, c1 = c10 * (1 - ty) + c11 * ty;
  double v = c0 * (1 - tz) + c1 * tz;
  if (grad) {
    double ddx0 = (c[0][0][1] - c[0][0][0]) * (1 - ty) + (c[0][1][1] - c[0][1][0]) * ty;
    double ddx1 = (c[1][0][1] - c[1][0][0]) * (1 - ty) + (c[1][1][1] - c[1][1][0]) * ty;
    grad[0] = (ddx0 * (1 - tz) + ddx1 * tz) * dsc[0];
    grad[1] = ((c01 - c00) * (1 - tz) + (c11 - c10) * tz) * dsc[1];
    grad[2] = (c1 - c0) * dsc[2];
  }
  return v;
}

static void dh_matrix_d(const double* p, double value, double m[12])
{
  double th = p[0], dd = p[1];
  if (p[4] != 0.0) dd += value; else th += value;
  double st = sin(th), ct = cos(th), sa = sin(p[3]), ca = cos(p[3]), a = p[2];
  m[0] = ct; m[1] = -st * ca; m[2] = st * sa; m[3] = a * ct;
  m[4] = st; m[5] = ct * ca; m[6] = -ct * sa; m[7] = a * st;
  m[8] = 0; m[9] = sa; m[10] = ca; m[11] = dd;
}

static void affine_mul_d(const double* x, const double* y, double* out)
{
  double r[12];
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) r[4 * i + j] = x[4 * i] * y[j] + x[4 * i + 1] * y[4 + j] + x[4 * i + 2] * y[8 + j];
    r[4 * i + 3] = x[4 * i] * y[3] + x[4 * i + 1] * y[7] + x[4 * i + 2] * y[11] + x[4 * i + 3];
  }
  memcpy(out, r, sizeof(r));
}

void gvo_fk_chain_d(const double* dh, uint32_t nl, const double base[12], const double* q, double* out)
{
  double t[12], a[12];
  memcpy(t, base, sizeof(t));
  for (uint32_t l = 0; l < nl; ++l) {
    memcpy(out + 12 * l, t, sizeof(t));
    dh_matrix_d(dh + 5 * l, q[l], a);
    affine_mul_d(t, a, t);
  }
}

/* c = sum w * max(0, r + eps - d(x))^2 over the spheres of every waypoint of a
   trajectory; sphere centres x = T_l(q) * c_local.  Gradients by the chain
   rule: dc/dx = -2 w h grad d;  dc/dq_j = sum over links l > j of
   z_j . ((x - o_j) x dc/dx) for revolute joints (z_j . dc/dx for prismatic),
   with z_j, o_j the z axis and origin of the frame T_j. */
void gvo_cost_d(const float* g, const uint32_t d[3], double side, const double* dh, uint32_t nl,
                const double base[12], const double* sph, uint32_t spl, const double* q,
                uint32_t ntraj, uint32_t nwp, double eps, double w, double* cost_out,
                double* dq_out, double* dx_out)
{
  double* T = (double*)malloc(12 * (nl + 1) * sizeof(double));
  for (uint32_t tr = 0; tr < ntraj; ++tr) {
    double c = 0.0;
    for (uint32_t wp = 0; wp < nwp; ++wp) {
      const double* qq = q + ((uint64_t)tr * nwp + wp) * nl;
      gvo_fk_chain_d(dh, nl, base, qq, T);
      double dq[64];
      for (uint32_t j = 0; j < nl && j < 64; ++j) dq[j] = 0.0;
      for (uint32_t l = 0; l < nl; ++l) {
        const double* m = T + 12 * l;
        for (uint32_t s = 0; s < spl; ++s) {
          const double* sp = sph + 4 * ((uint64_t)l * spl + s);
          double x[3], gd[3];
          for (int r = 0; r < 3; ++r) x[r] = m[4 * r] * sp[0] + m[4 * r + 1] * sp[1] + m[4 * r + 2] * sp[2] + m[4 * r + 3];
          double dv = gvo_trilinear_d(g, d, side, x, gd);
          double h = sp[3] + eps - dv;
          double gx[3] = {0, 0, 0};
          if (h > 0) {
            c += w * h * h;
            for (int r = 0; r < 3; ++r) gx[r] = -2.0 * w * h * gd[r];
          }
          if (dx_out) {
            double* o = dx_out + 3 * ((((uint64_t)tr * nwp + wp) * nl + l) * spl + s);
            o[0] = gx[0]; o[1] = gx[1]; o[2] = gx[2];
          }
          if (h > 0) {
            for (uint32_t j = 0; j < l; ++j) {
              /* joint j sits between frame T_j and T_{j+1}; it rotates about the
                 z axis of T_j (standard DH: A_j = Rz(theta) Tz(d) Tx(a) Rx(alpha)) */
              const double* f = T + 12 * j;
              double z[3] = {f[2], f[6], f[10]}, o[3] = {f[3], f[7], f[11]};
              if (dh[5 * j + 4] != 0.0) {
                dq[j] += z[0] * gx[0] + z[1] * gx[1] + z[2] * gx[2];
              } else {
                double rx = x[0] - o[0], ry = x[1] - o[1], rz = x[2] - o[2];
                double cx = ry * gx[2] - rz * gx[1], cy = rz * gx[0] - rx * gx[2], cz = rx * gx[1] - ry * gx[0];
                dq[j] += z[0] * cx + z[1] * cy + z[2] * cz;
              }
            }
          }
        }
      }
      if (dq_out)
        for (uint32_t j = 0; j < nl; ++j) dq_out[((uint64_t)tr * nwp + wp) * nl + j] = dq[j];
    }
    cost_out[tr] = c;
  }
  free(T);
}

/* ---- sensor data ------------------------------------------------------ */
uint64_t gvo_ray_cells(const uint32_t from[3], const uint32_t to[3], uint32_t* out)
{
  int32_t d[3], inc[3];
  for (int a = 0; a < 3; ++a) {
    if (to[a] > from[a]) { d[a] = (int32_t)(to[a] - from[a]); inc[a] = 1; }
    else { d[a] = (int32_t)(from[a] - to[a]); inc[a] = -1; }
  }
  int32_t x = (int32_t)from[0], y = (int32_t)from[1], z = (int32_t)from[2];
  int64_t n = 1 + (int64_t)d[0] + d[1] + d[2];
  int32_t exy = d[0] - d[1], exz = d[0] - d[2], eyz = d[1] - d[2];
  const int32_t dx2 = 2 * d[0], dy2 = 2 * d[1], dz2 = 2 * d[2];
  uint64_t k = 0;
  for (; n > 0; --n, ++k) {
    if (out) { out[3 * k] = (uint32_t)x; out[3 * k + 1] = (uint32_t)y; out[3 * k + 2] = (uint32_t)z; }
    if (exy > 0 && exz > 0) { x += inc[0]; exy -= dy2; exz -= dz2; }
    else if (eyz > 0) { y += inc[1]; exy += dx2; eyz -= dz2; }
    else { z += inc[2]; exz += dx2; eyz += dy2; }
  }
  return k;
}

static int sensor_point_ok(const uint32_t d[3], float side, const uint32_t s[3], const float* p,
                           const uint8_t* robot, uint32_t bit, uint32_t c[3])
{
  if (p[0] != p[0] || p[1] != p[1] || p[2] != p[2]) return 0;
  for (int a = 0; a < 3; ++a) c[a] = gvo_quantize(p[a], side);
  if (!(c[0] < d[0] && c[1] < d[1] && c[2] < d[2] && s[0] < d[0] && s[1] < d[1] && s[2] < d[2])) return 0;
  if (robot) {
    const uint8_t* v = robot + 32 * (((uint64_t)c[2] * d[1] + c[1]) * d[0] + c[0]);
    if ((v[bit >> 3] >> (bit & 7)) & 1) return 0;
  }
  return 1;
}

void gvo_prob_insert_sensor(int8_t* map, const uint32_t d[3], float side, const float sensor[3],
                            const float* pts, uint64_t n, int raycast, const uint8_t* robot, uint32_t bit)
{
  const uint64_t nv = (uint64_t)d[0] * d[1] * d[2];
  if (n > nv) n = nv;
  uint32_t s[3] = {gvo_quantize(sensor[0], side), gvo_quantize(sensor[1], side), gvo_quantize(sensor[2], side)};
  uint32_t c[3];
  if (raycast) {
    for (uint64_t i = 0; i < n; ++i) {
      if (!sensor_point_ok(d, side, s, pts + 3 * i, robot, bit, c)) continue;
      uint64_t m = gvo_ray_cells(s, c, NULL);
      uint32_t* cells = (uint32_t*)malloc(m * 12);
      gvo_ray_cells(s, c, cells);
      for (uint64_t k = 0; k < m; ++k) {
        uint64_t id = ((uint64_t)cells[3 * k + 2] * d[1] + cells[3 * k + 1]) * d[0] + cells[3 * k];
        map[id] = gvo_prob_update(map[id], -10); /* cSENSOR_MODEL_FREE, common_defines.h:159 */
      }
      free(cells);
    }
  }
  for (uint64_t i = 0; i < n; ++i) {
    if (!sensor_point_ok(d, side, s, pts + 3 * i, robot, bit, c)) continue;
    uint64_t id = ((uint64_t)c[2] * d[1] + c[1]) * d[0] + c[0];
    map[id] = gvo_prob_update(map[id], 72); /* cSENSOR_MODEL_OCCUPIED, common_defines.h:160 */
  }
}

/* ---- swept volumes ----------------------------------------------------- */
/* performLeftShift (helpers/BitVector.h:361-400), restated literally: a 64-bit
   buffer slides over the bytes; byte k of the result = (buffer >> shift) & 0xFF
   with the buffer holding bytes k .. k+7; byte 0 keeps only bits 4..7. */
void gvo_perform_left_shift(uint8_t v[32], uint32_t shift)
{
  uint64_t buffer = 0;
  for (uint32_t b = 0; b < 8; ++b) buffer += (uint64_t)v[b] << (b * 8);
  for (uint32_t b = 0; b < 32; ++b) {
    uint8_t nb = (uint8_t)(buffer >> shift);
    if (b == 0) nb &= 0xF0;
    v[b] = nb;
    buffer >>= 8;
    if (b + 8 < 32) buffer += (uint64_t)v[b + 8] << 56;
  }
}

/* shiftLeftSweptVolumeIDs on every voxel (BitVoxelMap.hpp:351-362 ->
   kernelShiftBitVector VoxelMapOperations.hpp:815-823; BitVoxelList.hpp:641-664);
   shifts above 56 are refused (returns 0) */
int gvo_shift_sv(uint8_t* voxels, uint64_t n, uint32_t shift)
{
  if (shift > 56) return 0;
  for (uint64_t i = 0; i < n; ++i) gvo_perform_left_shift(voxels + i * BV_BYTES, shift);
  return 1;
}

/* BitVoxelList::collideWithBitcheck (voxellist/BitVoxelList.hpp:260-292): ids
   present in both lists whose voxels collide: margin 0 = any common bit of the
   full vectors (BitvectorCollision, BitVoxelList.h:40-48), else
   bitMarginCollisionCheck(a, b, margin, 0) (BitVoxelList.h:56-73). */
uint64_t gvo_list_collide_bitcheck(const uint32_t* ia, const uint8_t* ba, uint64_t na, const uint32_t* ib,
                                   const uint8_t* bb, uint64_t nb, uint32_t margin)
{
  uint64_t cnt = 0, p;
  for (uint64_t i = 0; i < na; ++i) {
    if (!bsearch_u32(ib, nb, ia[i], &p)) continue;
    const uint8_t* va = ba + i * BV_BYTES;
    const uint8_t* vb = bb + p * BV_BYTES;
    int hit = 0;
    if (margin == 0) {
      for (int k = 0; k < 32; ++k) hit |= (va[k] & vb[k]) != 0;
    } else {
      uint8_t tmp[32];
      memset(tmp, 0, 32);
      hit = gvo_bitmargin_collision(va, vb, tmp, margin, 0);
    }
    cnt += hit ? 1 : 0;
  }
  return cnt;
}

/* BitVoxelList::collideCountingPerMeaning with a BitVectorVoxelList
   (BitVoxelList.hpp:295-343): for every id of this list that the other list
   also holds, every set bit of this list's voxel adds one to its meaning's
   counter; returns the sum of the counters. */
uint64_t gvo_list_collide_per_meaning(const uint32_t* ia, const uint8_t* ba, uint64_t na, const uint32_t* ib,
                                      uint64_t nb, uint64_t counts[256])
{
  uint64_t total = 0, p;
  memset(counts, 0, 256 * sizeof(uint64_t));
  for (uint64_t i = 0; i < na; ++i) {
    if (!bsearch_u32(ib, nb, ia[i], &p)) continue;
    for (uint32_t j = 0; j < 256; ++j)
      if ((ba[i * BV_BYTES + (j >> 3)] >> (j & 7)) & 1) {
        ++counts[j];
        ++total;
      }
  }
  return total;
}
