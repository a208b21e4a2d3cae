/*
 * gvx_oracle.h -- CPU restatement of GPU-Voxels' data-parallel core.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product path never links, loads or calls it.
 *
 * Every function restates the reference semantics serially and cites the
 * reference file:line it follows (paths relative to packages/gpu_voxels/src/
 * gpu_voxels/).  Where the reference is racy, lossy or undefined the oracle
 * states the rule it uses instead (see the comment at each function).
 *
 * Voxel layouts (identical to the reference, frozen):
 *   BitVectorVoxel     32 bytes; bit i in byte i>>3, position i&7
 *                      (helpers/BitVector.h:41-352)
 *   ProbabilisticVoxel int8 log-odds, UNKNOWN = -128 (voxel/ProbabilisticVoxel.h)
 *   DistanceVoxel      uint32 = x | y<<10 | z<<20 (voxel/DistanceVoxel.hpp:31-36)
 * Linear voxel index = z*dx*dy + y*dx + x (voxelmap/kernels/VoxelMapOperations.h:44-49)
 */
#ifndef GVX_ORACLE_H
#define GVX_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* map type numbering follows helpers/common_defines.h:66-79 (MapType) */
enum { GVO_BITVECTOR_MAP = 0, GVO_PROB_MAP = 4, GVO_DIST_MAP = 9 };

/* ---- scalar helpers ---------------------------------------------------- */
uint32_t gvo_quantize(float p, float side);
void gvo_sincosf(float x, float* s, float* c);
void gvo_dh_matrix(float theta, float d, float a, float alpha, int prismatic, float value, float m[12]);
void gvo_affine_mul(const float x[12], const float y[12], float out[12]);
void gvo_transform_points(const float m[12], const float* pts, uint64_t n, float* out);
int8_t gvo_float_to_probability(float v);
int8_t gvo_prob_update(int8_t cur, int8_t delta);
int8_t gvo_prob_insert(int8_t cur, uint32_t meaning);
int gvo_bitmargin_collision(const uint8_t v1[32], const uint8_t v2[32], uint8_t out[32],
                            uint32_t margin, uint32_t sv_offset);
uint32_t gvo_dist_uninit(const uint32_t dims[3]);
int32_t gvo_dist_sq(uint32_t raw, int32_t x, int32_t y, int32_t z, uint32_t uninit);

/* ---- dense maps ------------------------------------------------------ */
void gvo_map_clear(int type, void* map, const uint32_t dims[3]);
void gvo_map_insert_points(int type, void* map, const uint32_t dims[3], float side,
                           const float* pts, uint64_t n, uint32_t meaning, int* outside);
/* the same over `threads` OpenMP threads (same map) */
void gvo_map_insert_points_mt(int type, void* map, const uint32_t dims[3], float side, const float* pts,
                              uint64_t n, uint32_t meaning, int* outside, int threads);
void gvo_map_insert_coords(int type, void* map, const uint32_t dims[3],
                           const uint32_t* coords, uint64_t n, uint32_t meaning, int* outside);
void gvo_map_insert_meta(int type, void* map, const uint32_t dims[3], float side,
                         const float* pts, const uint32_t* cloud_sizes, uint32_t n_clouds,
                         const uint32_t* meanings, uint32_t meaning, int* outside);
int gvo_map_insert_meta_selfcoll(uint8_t* map, const uint32_t dims[3], float side, const float* pts,
                                 const uint32_t* cloud_sizes, uint32_t n_clouds, const uint32_t* meanings,
                                 const uint8_t* masks, uint8_t colliding[32], int* outside);
uint64_t gvo_map_collide(int type_a, void* a, int type_b, const void* b, const uint32_t dims[3],
                         float threshold, const int32_t offset[3]);
/* the same over `threads` OpenMP threads (same count and marks) */
uint64_t gvo_map_collide_mt(int type_a, void* a, int type_b, const void* b, const uint32_t dims[3], float thr,
                            const int32_t offset[3], int threads);
uint64_t gvo_map_collide_types(int type_a, void* a, int type_b, const void* b,
                               const uint32_t dims[3], float threshold, uint8_t mask[32]);
void gvo_map_clear_bits(void* map, const uint32_t dims[3], const uint8_t bits[32]);
uint64_t gvo_map_collide_sv(int type_a, void* a, int type_b, const void* b, const uint32_t dims[3],
                            float threshold, uint32_t window, uint32_t sv_offset, uint8_t mask[32]);
void gvo_perform_left_shift(uint8_t v[32], uint32_t shift);
int gvo_shift_sv(uint8_t* voxels, uint64_t n, uint32_t shift);
uint64_t gvo_list_collide_bitcheck(const uint32_t* ia, const uint8_t* ba, uint64_t na, const uint32_t* ib,
                                   const uint8_t* bb, uint64_t nb, uint32_t margin);
uint64_t gvo_list_collide_per_meaning(const uint32_t* ia, const uint8_t* ba, uint64_t na, const uint32_t* ib,
                                      uint64_t nb, uint64_t counts[256]);
/* RayCaster::rayCast (voxelmap/kernels/VoxelMapOperations.h:218-340): the 3-D
   integer walk from `from` to `to`.  Writes the visited cells (3 x u32 each) to
   out when it is not NULL; returns their number, 1 + |dx| + |dy| + |dz|. */
uint64_t gvo_ray_cells(const uint32_t from[3], const uint32_t to[3], uint32_t* out);
/* ProbVoxelMap::insertSensorData (voxelmap/ProbVoxelMap.hpp:51-78) ->
   kernelInsertSensorData (VoxelMapOperations.hpp:763-809) with one scan made
   deterministic: the FREE update of every cell on every ray first, then the
   OCCUPIED update of every measured voxel.  Points: the first min(n, voxels)
   ones, NaN skipped, point and sensor voxel both inside the map; with a robot
   map (32-byte voxels, same dims) a point whose voxel carries bit robot_bit is
   skipped entirely. */
void gvo_prob_insert_sensor(int8_t* map, const uint32_t dims[3], float side, const float sensor[3],
                            const float* pts, uint64_t n, int raycast, const uint8_t* robot, uint32_t robot_bit);

/* ---- voxel lists ----------------------------------------------------- */
/* A list is three caller-owned arrays of capacity `cap`: ids (u32), coords (3 x u32),
   bits (32 bytes each).  `*len` is the current length. */
void gvo_list_insert_points(uint32_t* ids, uint32_t* coords, uint8_t* bits, uint64_t* len,
                            const uint32_t dims[3], float side, const float* pts, uint64_t n,
                            const uint32_t* cloud_sizes, uint32_t n_clouds,
                            const uint32_t* meanings, uint32_t meaning);
void gvo_list_insert_coords(uint32_t* ids, uint32_t* coords, uint8_t* bits, uint64_t* len,
                            const uint32_t dims[3], const uint32_t* in, uint64_t n, uint32_t meaning);
void gvo_list_make_unique(uint32_t* ids, uint32_t* coords, uint8_t* bits, uint64_t* len);
void gvo_list_append_list(uint32_t* ids, uint32_t* coords, uint8_t* bits, uint64_t* len,
                          const uint32_t dims[3], const uint32_t* oids, const uint32_t* ocoords,
                          const uint8_t* obits, uint64_t on, const int32_t offset[3],
                          int has_meaning, uint32_t meaning);
uint64_t gvo_list_append_map(uint32_t* ids, uint32_t* coords, uint8_t* bits, uint64_t* len,
                             const uint32_t dims[3], const uint8_t* bitmap,
                             const int32_t offset[3], int has_meaning, uint32_t meaning);
uint64_t gvo_list_collide_map(uint32_t* ids, uint8_t* bits, uint64_t len, int map_type,
                              const void* map, const uint32_t dims[3], float threshold,
                              const int32_t offset[3], int mode, const uint8_t typemask[32],
                              uint8_t mask_out[32]);
uint64_t gvo_list_collide_list(const uint32_t* ids_a, const uint8_t* bits_a, uint64_t na,
                               const uint32_t* ids_b, const uint8_t* bits_b, uint64_t nb,
                               uint8_t mask_out[32]);

/* ---- distance maps --------------------------------------------------- */
void gvo_dist_merge_occupied(uint32_t* dist, const int8_t* prob, const uint32_t dims[3],
                             const uint32_t offset[3], float threshold);
void gvo_pba(uint32_t* map, const uint32_t dims[3]);
void gvo_pba_mt(uint32_t* map, const uint32_t dims[3], int threads);
void gvo_edt_bruteforce(const uint32_t* map, const uint32_t dims[3], uint32_t* out);
void gvo_exact_distances(uint32_t* map, const uint32_t dims[3], float side,
                         const float* obstacles, uint64_t n_obst);
void gvo_dist_squared(const uint32_t* map, const uint32_t dims[3], const uint32_t* idx,
                      uint64_t n, int32_t* out);

/* ---- robot / batched collision ------------------------------------------ */
void gvo_fk_chain(const float* dh /* n_links x 5: theta,d,a,alpha,prismatic */, uint32_t n_links,
                  const float base[12], const float* q, float* out /* n_links x 12 */);
void gvo_batch_collide(const uint8_t* env_bitmap /* 1 bit per voxel, linear index */,
                       const uint32_t dims[3], float side, const float* link_pts,
                       const uint32_t* link_sizes, uint32_t n_links, const float* dh,
                       const float base[12], const float* q, uint32_t n_poses,
                       uint32_t* counts, uint8_t* masks);
/* the same over `threads` OpenMP threads (all-cores CPU baseline) */
void gvo_batch_collide_mt(const uint8_t* env_bitmap /* 1 bit per voxel, linear index */,
                       const uint32_t dims[3], float side, const float* link_pts,
                       const uint32_t* link_sizes, uint32_t n_links, const float* dh,
                       const float base[12], const float* q, uint32_t n_poses,
                       uint32_t* counts, uint8_t* masks, int threads);

/* ---- differentiable clearance cost ---------------------------------------- */
double gvo_trilinear_d(const float* grid, const uint32_t dims[3], double side,
                       const double x[3], double grad[3]);
void gvo_fk_chain_d(const double* dh, uint32_t n_links, const double base[12],
                    const double* q, double* out);
void gvo_cost_d(const float* grid, const uint32_t dims[3], double side,
                const double* dh, uint32_t n_links, const double base[12],
                const double* spheres /* n_links*spl x 4 : cx,cy,cz,r (link frame) */,
                uint32_t spl, const double* q, uint32_t n_traj, uint32_t n_wp,
                double eps, double w, double* cost_out, double* dq_out, double* dx_out);

#ifdef __cplusplus
}
#endif
#endif
