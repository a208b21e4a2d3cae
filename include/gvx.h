/*
 * gvx.h -- C ABI of the MI355X-native GPU-Voxels core (libgvx.so).
 *
 * Plain pointers and sizes only.  Every entry point returns a gvx_status; on
 * failure gvx_last_error() holds a message for the calling thread.  Calls are
 * synchronous (they return after the GPU work has finished), like the
 * reference's API, unless stated otherwise.  All GPU work runs on one HIP
 * stream owned by the library (gvx_get_stream).
 *
 * Each declaration cites the reference interface it replaces; paths are
 * relative to packages/gpu_voxels/src/gpu_voxels/ of GPU-Voxels.
 *
 * Layouts (frozen, identical to the reference):
 *   BitVectorVoxel      32 bytes, bit i at byte i>>3 position i&7   (helpers/BitVector.h:41-352)
 *   ProbabilisticVoxel  int8 log-odds, -128 = unknown              (voxel/ProbabilisticVoxel.h:34-132)
 *   DistanceVoxel       uint32 x | y<<10 | z<<20                    (voxel/DistanceVoxel.hpp:31-65)
 *   voxel index         z*dx*dy + y*dx + x                          (voxelmap/kernels/VoxelMapOperations.h:44-49)
 *   points              float32 xyz triples (Vector3f, helpers/cuda_vectors.h:42-56)
 *   coordinates         uint32 xyz triples (Vector3ui)
 */
#ifndef GVX_H
#define GVX_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  GVX_OK = 0,
  GVX_ERR_BAD_ARG = 1,
  GVX_ERR_OOM = 2,
  GVX_ERR_DIM_MISMATCH = 3,
  GVX_ERR_UNSUPPORTED = 4,
  GVX_ERR_HIP = 5,
  GVX_ERR_NO_DEVICE = 6
} gvx_status;

/* MapType numbering of helpers/common_defines.h:66-79 */
typedef enum {
  GVX_MT_BITVECTOR_VOXELMAP = 0,
  GVX_MT_BITVECTOR_VOXELLIST = 1,
  GVX_MT_PROBAB_VOXELMAP = 4,
  GVX_MT_DISTANCE_VOXELMAP = 9
} gvx_map_type;

/* where a caller's buffer lives */
typedef enum { GVX_MEM_HOST = 0, GVX_MEM_DEVICE = 1 } gvx_mem;

typedef struct gvx_map_s* gvx_map;
typedef struct gvx_list_s* gvx_list;
typedef struct gvx_batch_s* gvx_batch;
typedef struct gvx_cost_s* gvx_cost;

/* ---- library ---------------------------------------------------------- */
/* cuTestAndInitDevice (helpers/cuda_handling.cu:102-130): select a device, create the stream. */
int gvx_init(int device);
const char* gvx_last_error(void);
/* the library's HIP stream (hipStream_t) */
int gvx_get_stream(void** stream);
/* per-kernel HIP-event timing on the library stream (replaces the icl_core
   PerformanceMonitor, icl_core_performance_monitor/PerformanceMonitorMacros.h:43-69) */
int gvx_profile_enable(int enable);
int gvx_profile_get(const char* kernel, double* avg_ms, uint64_t* launches);
int gvx_profile_reset(void);

/* ---- dense maps ---------------------------------------------------------- */
/* TemplateVoxelMap ctor (voxelmap/TemplateVoxelMap.hpp:62-144); 64-bit sizes, so
   512^3 BitVector and 1024^3 Distance maps are allowed (the reference refuses
   >= 2^32 bytes).  Distance maps with an axis of 1024 use 0xFFFFFFFF as the
   "uninitialised" voxel (the reference's 1023 sentinel is a valid coordinate there). */
int gvx_map_create(int type, uint32_t dx, uint32_t dy, uint32_t dz, float side, gvx_map* out);
int gvx_map_destroy(gvx_map map);
int gvx_map_info(gvx_map map, int* type, uint32_t dims[3], float* side, uint64_t* bytes);
/* getDeviceDataPtr (TemplateVoxelMap.h:69-87): zero-copy access to the voxels */
int gvx_map_device_ptr(gvx_map map, void** ptr);
/* clearMap (TemplateVoxelMap.hpp:197-271) */
int gvx_map_clear(gvx_map map);
/* BitVoxelMap::clearBit / clearBits (voxelmap/BitVoxelMap.hpp:62-85) */
int gvx_map_clear_bits(gvx_map map, const uint8_t bits[32]);
/* copy the whole voxel array out / in (GVX_MEM_HOST or GVX_MEM_DEVICE buffers) */
int gvx_map_read(gvx_map map, void* dst, int mem);
int gvx_map_write(gvx_map map, const void* src, int mem);
/* insertPointCloud -> kernelInsertGlobalPointCloud (TemplateVoxelMap.hpp:666-709,
   VoxelMapOperations.hpp:241-300).  *points_outside is set when a point fell outside. */
int gvx_map_insert_points(gvx_map map, const float* xyz, uint64_t n, uint32_t meaning, int mem,
                          int* points_outside);
/* insertCoordinateList -> kernelInsertCoordinateTuples (TemplateVoxelMap.hpp:711-746,
   VoxelMapOperations.hpp:302-351) */
int gvx_map_insert_coords(gvx_map map, const uint32_t* xyz, uint64_t n, uint32_t meaning, int mem,
                          int* points_outside);
/* insertMetaPointCloud (TemplateVoxelMap.hpp:850-904, VoxelMapOperations.hpp:440-575):
   n_clouds consecutive clouds in xyz; meanings[c] per cloud, or `meaning` for all when NULL. */
int gvx_map_insert_meta(gvx_map map, const float* xyz, const uint32_t* cloud_sizes, uint32_t n_clouds,
                        const uint32_t* meanings, uint32_t meaning, int mem, int* points_outside);
/* BitVoxelMap::insertMetaPointCloudWithSelfCollisionCheck (voxelmap/BitVoxelMap.hpp:228-341,
   kernelInsertMetaPointCloudSelfCollCheck VoxelMapOperations.hpp:582-638); map must be a
   BitVector map.  Sub-clouds are applied in order; within one, every point tests its voxel
   as it was before the sub-cloud: masked = voxel & masks[c] holding a bit other than
   eBVM_FREE makes the point collide (colliding |= masked, colliding gets meanings[c], the
   voxel gets eBVM_COLLISION); then every point sets meanings[c].  meanings NULL:
   eBVM_SWEPT_VOLUME_START + c; masks (n_clouds x 32 bytes) NULL: every bit but
   eBVM_COLLISION and the cloud's own meaning.  colliding (32 bytes, may be NULL), and
   *self_collision = colliding != 0. */
int gvx_map_insert_meta_selfcoll(gvx_map map, const float* xyz, const uint32_t* cloud_sizes, uint32_t n_clouds,
                                 const uint32_t* meanings, const uint8_t* masks, int mem, uint8_t colliding[32],
                                 int* self_collision, int* points_outside);
/* ProbVoxelMap::insertSensorData (voxelmap/ProbVoxelMap.hpp:51-78 ->
   kernelInsertSensorData, VoxelMapOperations.hpp:763-809, RayCaster
   VoxelMapOperations.h:218-340): map must be a ProbVoxelMap.  For the first
   min(n, voxels) points that are not NaN and whose voxel and the sensor's voxel lie
   in the map: with raycasting, every cell on the integer line from the sensor voxel
   to the point voxel gets the FREE update (-10); then every point voxel gets the
   OCCUPIED update (+72).  One call is deterministic: all FREE updates first, then
   all OCCUPIED ones (the reference interleaves them racily).  robot_bits_dev: NULL,
   or a device array of 32-byte BitVector voxels of the same dims (cut_real_robot):
   points whose voxel carries bit robot_meaning there are skipped. */
int gvx_map_insert_sensor(gvx_map map, const float* xyz, uint64_t n, const float sensor_pose[3],
                          int enable_raycasting, const void* robot_bits_dev, uint32_t robot_meaning, int mem);
/* collideWith (BitVoxelMap.hpp:199-210, ProbVoxelMap.hpp:98-109 ->
   TemplateVoxelMap.hpp:590-621, VoxelMapOperations.hpp:129-184): DefaultCollider
   count; voxel this[i + offset] is compared with other[i]; colliding voxels of
   `self` get insert(eBVM_COLLISION).  offset may be NULL. */
int gvx_map_collide(gvx_map self, gvx_map other, float threshold, const int32_t offset[3],
                    uint64_t* count);
/* collideWithTypes (BitVoxelMap.hpp:92-224, VoxelMapOperations.hpp:187-239):
   SVCollider count + OR of the colliding meanings (self must be a BitVector map). */
int gvx_map_collide_types(gvx_map self, gvx_map other, float threshold, uint64_t* count,
                          uint8_t mask[32]);

/* collisionCheckBitvector with SVCollider(threshold, window) and sv_offset
   (voxelmap/BitVoxelMap.hpp:90-195, kernelCollideVoxelMapsBitvector
   VoxelMapOperations.hpp:186-239, SVCollider.hpp:40-110): Bit x Bit voxels collide
   by bitMarginCollisionCheck (helpers/BitVector.h:415-471) with margin = window;
   Bit x Prob ignores window and sv_offset (as collideWithTypes).  Colliding voxels
   of self get eBVM_COLLISION; mask = OR of every colliding voxel's result.
   window <= 31 and sv_offset / 8 <= 31 - window (every shift of the check < 64). */
int gvx_map_collide_sv(gvx_map self, gvx_map other, float threshold, uint32_t window, uint32_t sv_offset,
                       uint64_t* count, uint8_t mask[32]);
/* shiftLeftSweptVolumeIDs (BitVoxelMap.hpp:351-362, kernelShiftBitVector
   VoxelMapOperations.hpp:815-823): performLeftShift (helpers/BitVector.h:361-400)
   on every voxel of a BitVector map; shift <= 56 */
int gvx_map_shift_sv(gvx_map map, uint32_t shift);

/* ---- distance maps ------------------------------------------------------ */
/* mergeOccupied (voxelmap/DistanceVoxelMap.hpp:109-136) */
int gvx_dist_merge_occupied(gvx_map dist, gvx_map prob, const uint32_t offset[3], float threshold);
/* parallelBanding3D (DistanceVoxelMap.hpp:284-645): exact EDT in place.  The
   banding arguments m1..m3 of the reference only change its schedule; any
   values are accepted and ignored. */
int gvx_dist_pba(gvx_map dist);
/* No counterpart in the reference.  The same exact EDT as gvx_dist_pba, for
   callers that cannot allocate device memory (tight planning loops, harnesses
   that forbid hipMalloc): it calls no hipMalloc/hipFree and works in a fixed
   32 MiB device workspace that the library holds on its device from gvx_init
   to the end of the process (12 B per voxel + 384 KiB + 4 B per (y, z) line + flag words,
   each rounded up to 256 B: 128^3 fits).  A larger map is rejected with
   GVX_ERR_BAD_ARG.  Every call on the device shares the one workspace; the
   library serialises its own calls, so only work the caller overlaps with a
   call outside the library's stream (a replayed capture, another stream)
   races on it. */
int gvx_dist_pba_noalloc(gvx_map dist);
/* exactDistances3D (DistanceVoxelMap.hpp:208-253, VoxelMapOperations.hpp:896-990) */
int gvx_dist_exact(gvx_map dist, const float* obstacles, uint64_t n, int mem);
/* getSquaredDistances / getSquaredDistancesToHost (DistanceVoxelMap.hpp:716-752);
   idx == NULL means every voxel in index order */
int gvx_dist_squared(gvx_map dist, const uint32_t* idx, uint64_t n, int32_t* out, int mem);
/* float metric distance per voxel, sqrt(squared) * side (input of the clearance cost) */
int gvx_dist_metric(gvx_map dist, float* out, int mem);

/* ---- voxel lists (BitVectorVoxelList) ------------------------------------------ */
/* BitVoxelList ctor (voxellist/BitVoxelList.hpp:38-49) */
int gvx_list_create(uint32_t dx, uint32_t dy, uint32_t dz, float side, gvx_list* out);
int gvx_list_destroy(gvx_list list);
int gvx_list_size(gvx_list list, uint64_t* n);
/* reference map dimensions and voxel side length the list was created with
   (m_ref_map_dim, getVoxelSideLength: voxellist/TemplateVoxelList.h:113-116, 277) */
int gvx_list_info(gvx_list list, uint32_t dims[3], float* side);
/* clearMap */
int gvx_list_clear(gvx_list list);
/* copy out ids (u32), coords (3 x u32), bits (32 B); any pointer may be NULL */
int gvx_list_read(gvx_list list, uint32_t* ids, uint32_t* coords, uint8_t* bits, int mem);
/* device pointers of the three arrays (getDeviceIdPtr/CoordPtr/..., TemplateVoxelList.h:94-105) */
int gvx_list_device_ptrs(gvx_list list, uint32_t** ids, uint32_t** coords, uint8_t** bits);
/* insertPointCloud / insertMetaPointCloud (TemplateVoxelList.hpp:363-514): append,
   remove_out_of_bounds, then make_unique (sort by id, OR duplicates).
   cloud_sizes == NULL: one cloud with `meaning`. */
int gvx_list_insert_points(gvx_list list, const float* xyz, uint64_t n, const uint32_t* cloud_sizes,
                           uint32_t n_clouds, const uint32_t* meanings, uint32_t meaning, int mem);
/* insertCoordinateList (TemplateVoxelList.hpp:395-440) */
int gvx_list_insert_coords(gvx_list list, const uint32_t* xyz, uint64_t n, uint32_t meaning, int mem);
/* merge(list) (TemplateVoxelList.hpp:878-916); meaning < 0: keep bits */
int gvx_list_merge_list(gvx_list self, gvx_list other, const int32_t offset[3], int meaning);
/* merge(BitVectorVoxelMap) (TemplateVoxelList.hpp:969-1030) */
int gvx_list_merge_map(gvx_list self, gvx_map map, const int32_t offset[3], int meaning);
/* list x map: collideWith (mode 0, BitVoxelList.hpp:63-77 -> TemplateVoxelList.hpp:1170-1213),
   collideWithTypes (mode 1, BitVoxelList.hpp:120-211), collideWithTypeMask (mode 2,
   BitVoxelList.hpp:213-258); kernels VoxelListOperations.hpp:146-334.
   typemask used in mode 2, mask_out filled in mode 1 (either may be NULL otherwise). */
int gvx_list_collide_map(gvx_list self, gvx_map map, float threshold, const int32_t offset[3], int mode,
                         const uint8_t typemask[32], uint64_t* count, uint8_t mask_out[32]);
/* list x list: collideWith (BitVoxelList.hpp:79-92) and collideWithTypes
   (BitVoxelList.hpp:94-118, OR of both lists' bits at common voxels) */
int gvx_list_collide_list(gvx_list self, gvx_list other, uint64_t* count, uint8_t mask_out[32]);

/* ---- DH robot (robot/dh_robot/KinematicChain.cu:93-126, KinematicLink.cu:24-79) ------ */
/* dh: n_links x 5 floats (theta, d, a, alpha, prismatic); base: 3x4 row-major or NULL;
   q: n_links joint values.  Link l is posed by base * A_0(q_0) ... A_{l-1}(q_{l-1}). */
/* DHParameters::convertDHtoM (KinematicLink.cu:24-79) on the host, with the same
   deterministic sine/cosine as the GPU kernels; dh = (theta, d, a, alpha, prismatic) */
int gvx_dh_matrix(const float dh[5], float value, float out[12]);
/* setConfiguration + getTransformedClouds (MetaPointCloud::transformSubCloud,
   helpers/MetaPointCloud.cu:636-661): the posed link clouds into `out` (n x 3 floats) */
int gvx_robot_transform(const float* link_pts, const uint32_t* link_sizes, uint32_t n_links, const float* dh,
                        const float base[12], const float* q, int mem, float* out, int out_mem);
/* setConfiguration + insertRobotIntoMap (GpuVoxels.cpp:499-517) fused into one FK thread and
   one insert launch; meanings[l] per link or `meaning` for all when NULL */
int gvx_map_insert_robot(gvx_map map, const float* link_pts, const uint32_t* link_sizes, uint32_t n_links,
                         const float* dh, const float base[12], const float* q, const uint32_t* meanings,
                         uint32_t meaning, int mem, int* points_outside);
/* GpuVoxels::insertRobotIntoMapSelfCollAware (GpuVoxels.cpp:519-541): setConfiguration,
   then gvx_map_insert_meta_selfcoll of the posed link clouds */
int gvx_map_insert_robot_selfcoll(gvx_map map, const float* link_pts, const uint32_t* link_sizes, uint32_t n_links,
                                  const float* dh, const float base[12], const float* q, const uint32_t* meanings,
                                  const uint8_t* masks, int mem, uint8_t colliding[32], int* self_collision,
                                  int* points_outside);
/* the same into a voxel list (insertMetaPointCloud of the posed clouds) */
int gvx_list_insert_robot(gvx_list list, const float* link_pts, const uint32_t* link_sizes, uint32_t n_links,
                          const float* dh, const float base[12], const float* q, const uint32_t* meanings,
                          uint32_t meaning, int mem);

/* ---- batched multi-pose collision (no reference counterpart: it replaces
   the per-pose clearMap + setConfiguration + insertMetaPointCloud +
   collideWith loop of gvl_ompl_planning/gvl_ompl_planner_helper.cpp:169-200) ---- */
/* env: BitVector or Prob map (occupied = BitVoxel::isOccupied / prob >= floatToProbability(thr)).
   link_pts: link-frame points of all links, consecutive; dh: n_links x 5 floats
   (theta, d, a, alpha, prismatic) of robot/dh_robot/KinematicLink.cu:24-79; base: 3x4 row-major.
   Link l's points get meaning eBVM_SWEPT_VOLUME_START + l and are posed by
   base * A_0(q_0) ... A_{l-1}(q_{l-1}) (KinematicChain.cu:93-126). */
int gvx_batch_create(gvx_map env, float threshold, const float* link_pts, const uint32_t* link_sizes,
                     uint32_t n_links, const float* dh, const float base[12], uint32_t max_poses,
                     gvx_batch* out);
int gvx_batch_destroy(gvx_batch b);
/* rebuild the occupancy bitmap after the environment map changed */
int gvx_batch_refresh_env(gvx_batch b);
/* q: n_poses x n_links joint values.  counts[p]: distinct robot voxels in
   collision; masks[p]: 32-byte OR of the colliding link meanings. */
int gvx_batch_collide(gvx_batch b, const float* q, uint32_t n_poses, int mem, uint32_t* counts,
                      uint8_t* masks, int out_mem);
/* the same with device buffers only (4-byte aligned), enqueued on the library
   stream without waiting: calls run in order; gvx_sync (or the stream) waits */
int gvx_batch_collide_async(gvx_batch b, const float* q_dev, uint32_t n_poses, uint32_t* counts_dev,
                            uint8_t* masks_dev);
/* broad phase of the last batched call (waits for it): (chunk, pose) pairs
   tested, pairs that survived the bounding-box test, and their point-poses */
int gvx_batch_last_stats(gvx_batch b, uint64_t* tests, uint64_t* live_pairs, uint64_t* live_point_poses);
/* wait for all work enqueued on the library stream */
int gvx_sync(void);
/* choose and cache the fastest launch schedule for n_poses, timed on the joint values of
   the last gvx_batch_collide call with a host-memory q.  Schedules 0-7: chunk size,
   block size, poses per thread and one-block-per-item or persistent grid, launched
   directly; 8-15: the same kernels replayed from a HIP graph captured per
   (n_poses, buffers). */
int gvx_batch_autotune(gvx_batch b, uint32_t n_poses, int* schedule_out);
int gvx_batch_set_schedule(gvx_batch b, int schedule);
/* per-pose reference path on the GPU: clear a robot map, transform and insert
   every link, collideWith(env).  robot must be a BitVector map of env's size. */
int gvx_batch_collide_per_pose(gvx_batch b, gvx_map robot, const float* q, uint32_t n_poses, int mem,
                               uint32_t* counts);

/* ---- differentiable clearance cost (no reference counterpart) --------------- */
/* grid: float metric distances (gvx_dist_metric), device pointer, dims, side.
   spheres: n_links x spheres_per_link x 4 floats (cx, cy, cz, r) in link frames. */
int gvx_cost_create(const float* grid_dev, const uint32_t dims[3], float side, const float* dh,
                    uint32_t n_links, const float base[12], const float* spheres, uint32_t spheres_per_link,
                    float eps, float weight, gvx_cost* out);
int gvx_cost_destroy(gvx_cost c);
/* q: device pointer n_traj x n_wp x n_links; cost_out: device n_traj */
int gvx_cost_forward(gvx_cost c, const float* q_dev, uint32_t n_traj, uint32_t n_wp, float* cost_out_dev);
/* dq_out: device n_traj x n_wp x n_links; dx_out (may be NULL): device
   n_traj x n_wp x n_links x spheres_per_link x 3; upstream: device n_traj (dL/dc) or NULL for ones */
int gvx_cost_backward(gvx_cost c, const float* q_dev, uint32_t n_traj, uint32_t n_wp,
                      const float* upstream_dev, float* dq_out_dev, float* dx_out_dev);

/* BitVoxelList::shiftLeftSweptVolumeIDs (voxellist/BitVoxelList.hpp:641-664) */
int gvx_list_shift_sv(gvx_list list, uint32_t shift);
/* BitVoxelList::collideWithBitcheck (BitVoxelList.hpp:260-292): number of ids in
   both lists whose voxels collide: margin 0 = any common bit of the full vectors,
   else bitMarginCollisionCheck with sv_offset 0; margin <= 31 */
int gvx_list_collide_bitcheck(gvx_list self, gvx_list other, uint32_t margin, uint64_t* count);
/* BitVoxelList::collideCountingPerMeaning with a BitVectorVoxelList
   (BitVoxelList.hpp:295-343): counts[j] = number of self's voxels, at ids other
   also holds, with bit j set; *total (may be NULL) = the sum */
int gvx_list_collide_per_meaning(gvx_list self, gvx_list other, uint64_t counts[256], uint64_t* total);

/* ---- files: map and list dumps, point-cloud readers ---------------------------- */
/* TemplateVoxelMap::writeToDisk (voxelmap/TemplateVoxelMap.hpp:906-954): int32 map
   type, float side, 3 x u32 dims, then the raw voxels (all bytes; the reference's
   32-bit size counter truncates maps of 4 GiB and more). */
int gvx_map_write_file(gvx_map map, const char* path);
/* TemplateVoxelMap::readFromDisk (TemplateVoxelMap.hpp:956-1039): fails on a type
   or dimension mismatch; a different side length is accepted and reported through
   side_matches (may be NULL), as the reference only warns. */
int gvx_map_read_file(gvx_map map, const char* path, int* side_matches);
/* TemplateVoxelList::writeToDisk (voxellist/TemplateVoxelList.hpp:625-658): int32
   map type, 3 x u32 reference dims, float side, u32 count, ids, coords, voxels. */
int gvx_list_write_file(gvx_list list, const char* path);
/* TemplateVoxelList::readFromDisk (TemplateVoxelList.hpp:660-715): fails unless
   type, dims and side all match; the entries are taken as stored. */
int gvx_list_read_file(gvx_list list, const char* path);
/* PointcloudFileHandler::loadPointCloud (helpers/PointcloudFileHandler.cpp:69-137)
   with XyzFileReader.cpp:29-56, BinvoxFileReader.cpp:42-163 and the PCD reader
   (PcdFileReader.cpp:31-54, pcl::PointXYZ; ascii, binary and binary_compressed).
   The format is picked by the first of "xyz", "pcd", "binvox" found in the path;
   use_model_path prefixes $GPU_VOXELS_MODEL_PATH; then optional shift of the
   minimum to zero and p = scaling * p + offset.  *xyz is malloc'd (n x 3 floats),
   release it with gvx_free. */
int gvx_load_point_cloud(const char* path, int use_model_path, int shift_to_zero, const float offset[3],
                         float scaling, float** xyz, uint64_t* n);
void gvx_free(void* p);

#ifdef __cplusplus
}
#endif
#endif
