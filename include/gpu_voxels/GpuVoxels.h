// gpu_voxels/GpuVoxels.h -- C++ drop-in API of the MI355X-native GPU-Voxels core.
//
// Same namespaces, class names and method signatures as the reference's hot path
// (packages/gpu_voxels/src/gpu_voxels/: GpuVoxels.h, GpuVoxelsMap.h,
// helpers/CollisionInterfaces.h, voxelmap/*, voxellist/*, robot/dh_robot/*), with
// std::shared_ptr in place of boost::shared_ptr.  Everything is implemented over
// the C ABI of include/gvx.h (libgvx.so).  Return conventions follow the
// reference: bool for facade calls, SSIZE_MAX for a dimension mismatch in list
// collisions, `throw std::string` for map types that are not implemented.
#ifndef GPU_VOXELS_GPUVOXELS_H_INCLUDED
#define GPU_VOXELS_GPUVOXELS_H_INCLUDED

#include <climits>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <ostream>
#include <string>
#include <vector>

#include "../gvx.h"

namespace gpu_voxels {

// ---------------------------------------------------------------- common_defines.h
static const std::size_t BIT_VECTOR_LENGTH = 256;

enum BitVoxelMeaning {
  eBVM_FREE = 0,
  eBVM_OCCUPIED = 1,
  eBVM_COLLISION = 2,
  eBVM_UNKNOWN = 3,
  eBVM_SWEPT_VOLUME_START = 4,
  eBVM_SWEPT_VOLUME_END = 254,
  eBVM_MAX_FREE_PROB = 4,
  eBVM_UNCERTAIN_OCC_PROB = 129,
  eBVM_MAX_OCC_PROB = 254,
  eBVM_UNDEFINED = 255
};

enum MapType {
  MT_BITVECTOR_VOXELMAP,
  MT_BITVECTOR_VOXELLIST,
  MT_BITVECTOR_OCTREE,
  MT_BITVECTOR_MORTON_VOXELLIST,
  MT_PROBAB_VOXELMAP,
  MT_PROBAB_VOXELLIST,
  MT_PROBAB_OCTREE,
  MT_PROBAB_MORTON_VOXELLIST,
  MT_COUNTING_VOXELLIST,
  MT_DISTANCE_VOXELMAP
};

static const std::string GPU_VOXELS_MAP_TYPE_NOT_IMPLEMENTED = "THIS TYPE OF DATA STRUCTURE IS NOT YET IMPLEMENTED!";
static const std::string GPU_VOXELS_MAP_OPERATION_NOT_SUPPORTED = "THIS OPERATION IS NOT SUPPORTED BY THE DATA STRUCTURE!";

typedef int8_t Probability;
static const Probability UNKNOWN_PROBABILITY = Probability(-128);
static const Probability MIN_PROBABILITY = Probability(-127);
static const Probability MAX_PROBABILITY = Probability(127);
static const Probability cSENSOR_MODEL_FREE = -10;
static const Probability cSENSOR_MODEL_OCCUPIED = 72;
static const int32_t PBA_UNINITIALISED_COORD = 1023;
static const int32_t MAX_OBSTACLE_DISTANCE = 2147483647;

// ---------------------------------------------------------------- cuda_datatypes.h
struct Vector3f {
  float x, y, z;
  Vector3f() : x(0), y(0), z(0) {}
  explicit Vector3f(float v) : x(v), y(v), z(v) {}
  Vector3f(float a, float b, float c) : x(a), y(b), z(c) {}
  bool operator==(const Vector3f& o) const { return x == o.x && y == o.y && z == o.z; }
  bool operator!=(const Vector3f& o) const { return !(*this == o); }
  Vector3f operator+(const Vector3f& o) const { return Vector3f(x + o.x, y + o.y, z + o.z); }
  Vector3f operator-(const Vector3f& o) const { return Vector3f(x - o.x, y - o.y, z - o.z); }
};
struct Vector3ui {
  uint32_t x, y, z;
  Vector3ui() : x(0), y(0), z(0) {}
  explicit Vector3ui(uint32_t v) : x(v), y(v), z(v) {}
  Vector3ui(uint32_t a, uint32_t b, uint32_t c) : x(a), y(b), z(c) {}
  bool operator==(const Vector3ui& o) const { return x == o.x && y == o.y && z == o.z; }
  bool operator!=(const Vector3ui& o) const { return !(*this == o); }
  Vector3ui operator+(const Vector3ui& o) const { return Vector3ui(x + o.x, y + o.y, z + o.z); }
};
struct Vector3i {
  int32_t x, y, z;
  Vector3i() : x(0), y(0), z(0) {}
  explicit Vector3i(int32_t v) : x(v), y(v), z(v) {}
  Vector3i(int32_t a, int32_t b, int32_t c) : x(a), y(b), z(c) {}
  bool operator==(const Vector3i& o) const { return x == o.x && y == o.y && z == o.z; }
  bool operator!=(const Vector3i& o) const { return !(*this == o); }
};
struct Matrix4f {
  float a11, a12, a13, a14, a21, a22, a23, a24, a31, a32, a33, a34, a41, a42, a43, a44;
  static Matrix4f createIdentity()
  {
    Matrix4f m = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    return m;
  }
};

// ---------------------------------------------------------------- BitVector.h (host side)
template <std::size_t num_bits>
class BitVector {
 public:
  typedef uint8_t item_type;
  BitVector() { clear(); }
  void clear() { for (auto& b : m_bytes) b = 0; }
  bool getBit(uint32_t i) const { return (m_bytes[i >> 3] >> (i & 7)) & 1; }
  void setBit(uint32_t i) { m_bytes[i >> 3] |= item_type(1u << (i & 7)); }
  void clearBit(uint32_t i) { m_bytes[i >> 3] &= item_type(~(1u << (i & 7))); }
  item_type getByte(uint32_t i, uint8_t = 0) const { return m_bytes[i >> 3]; }
  void setByte(uint32_t i, item_type d) { m_bytes[i >> 3] = d; }
  bool isZero() const
  {
    for (auto b : m_bytes) if (b) return false;
    return true;
  }
  // true when no bit or only eBVM_FREE is set (BitVector.h:183-198)
  bool noneButEmpty() const
  {
    if (m_bytes[0] & 0xFE) return false;
    for (std::size_t i = 1; i < m_size; ++i) if (m_bytes[i]) return false;
    return true;
  }
  BitVector operator|(const BitVector& o) const { BitVector r; for (std::size_t i = 0; i < m_size; ++i) r.m_bytes[i] = m_bytes[i] | o.m_bytes[i]; return r; }
  BitVector operator&(const BitVector& o) const { BitVector r; for (std::size_t i = 0; i < m_size; ++i) r.m_bytes[i] = m_bytes[i] & o.m_bytes[i]; return r; }
  BitVector operator~() const { BitVector r; for (std::size_t i = 0; i < m_size; ++i) r.m_bytes[i] = item_type(~m_bytes[i]); return r; }
  void operator|=(const BitVector& o) { for (std::size_t i = 0; i < m_size; ++i) m_bytes[i] |= o.m_bytes[i]; }
  bool operator==(const BitVector& o) const { for (std::size_t i = 0; i < m_size; ++i) if (m_bytes[i] != o.m_bytes[i]) return false; return true; }
  const item_type* data() const { return m_bytes; }
  item_type* data() { return m_bytes; }
  // stream order: bit 0 first (BitVector.h:271-288)
  friend std::ostream& operator<<(std::ostream& os, const BitVector& dt)
  {
    for (std::size_t i = 0; i < num_bits; ++i) os << (dt.getBit((uint32_t)i) ? '1' : '0');
    return os;
  }

 protected:
  static const std::size_t m_size = (num_bits + 7) / 8;
  item_type m_bytes[m_size];
};

class BitVectorVoxel {
 public:
  BitVector<BIT_VECTOR_LENGTH>& bitVector() { return m_bit_vector; }
  const BitVector<BIT_VECTOR_LENGTH>& bitVector() const { return m_bit_vector; }
  void insert(BitVoxelMeaning m) { m_bit_vector.setBit(m); }
  bool isOccupied(float) const { return !m_bit_vector.noneButEmpty(); }
  friend std::ostream& operator<<(std::ostream& os, const BitVectorVoxel& v) { return os << v.m_bit_vector; }

 private:
  BitVector<BIT_VECTOR_LENGTH> m_bit_vector;
};

class ProbabilisticVoxel {
 public:
  // ProbabilisticVoxel.h:39-51
  static Probability floatToProbability(float val);
  static float probabilityToFloat(Probability val);
};

// ---------------------------------------------------------------- point clouds
class PointCloud {
 public:
  PointCloud() {}
  explicit PointCloud(const std::vector<Vector3f>& pts) : m_points(pts) {}
  void update(const std::vector<Vector3f>& pts) { m_points = pts; }
  uint32_t getPointCloudSize() const { return (uint32_t)m_points.size(); }
  const std::vector<Vector3f>& points() const { return m_points; }

 private:
  std::vector<Vector3f> m_points;
};

// MetaPointCloud (helpers/MetaPointCloud.h): several clouds in one contiguous buffer
class MetaPointCloud {
 public:
  MetaPointCloud() {}
  explicit MetaPointCloud(const std::vector<std::vector<Vector3f> >& clouds);
  MetaPointCloud(const std::vector<std::vector<Vector3f> >& clouds, const std::vector<std::string>& names);
  // MetaPointCloud.cu:116-171: one cloud per file (PointcloudFileHandler, no shift or offset),
  // named after the file unless names are given
  MetaPointCloud(const std::vector<std::string>& point_cloud_files, bool use_model_path);
  MetaPointCloud(const std::vector<std::string>& point_cloud_files, const std::vector<std::string>& point_cloud_names,
                 bool use_model_path);
  void addCloud(const std::vector<Vector3f>& cloud, const std::string& name = "");
  void updatePointCloud(uint16_t cloud, const std::vector<Vector3f>& pts);
  uint16_t getNumberOfPointclouds() const { return (uint16_t)m_sizes.size(); }
  uint32_t getAccumulatedPointcloudSize() const { return (uint32_t)m_points.size(); }
  std::vector<uint32_t> getPointcloudSizes() const { return m_sizes; }
  uint32_t getPointcloudSize(uint16_t cloud) const { return m_sizes[cloud]; }
  bool hasCloud(const std::string& name) const;
  int16_t getCloudNumber(const std::string& name) const;
  void syncToDevice() {}
  const std::vector<Vector3f>& points() const { return m_points; }

 private:
  std::vector<Vector3f> m_points;
  std::vector<uint32_t> m_sizes;
  std::vector<std::string> m_names;
};

namespace geometry_generation {
// GeometryGeneration.cpp:90-127
std::vector<Vector3f> createBoxOfPoints(Vector3f min, Vector3f max, float delta);
std::vector<Vector3ui> createBoxOfPoints(Vector3f min, Vector3f max, float delta, float voxel_side_length);
}  // namespace geometry_generation

namespace file_handling {
// helpers/PointcloudFileHandler.h: .xyz, .pcd (ascii / binary / binary_compressed) and
// .binvox files; the format is the first of "xyz", "pcd", "binvox" found in the path
class PointcloudFileHandler {
 public:
  static PointcloudFileHandler* Instance();
  bool loadPointCloud(const std::string path, const bool use_model_path, std::vector<Vector3f>& points,
                      const bool shift_to_zero = false, const Vector3f& offset_XYZ = Vector3f(),
                      const float scaling = 1.0);
  void centerPointCloud(std::vector<Vector3f>& points);
  void shiftPointCloudToZero(std::vector<Vector3f>& points);

 private:
  PointcloudFileHandler() {}
};
}  // namespace file_handling

// voxel/SVCollider.h: swept-volume collider; window = bit margin of the Bit x Bit check
class SVCollider {
 public:
  // the reference's default: both thresholds at probability 100 (227/254 maps to 100), window 10
  SVCollider() : m_threshold(227.0f / 254.0f), m_window(10) {}
  SVCollider(const float coll_threshold, const size_t window_size = 0)
      : m_threshold(coll_threshold), m_window(window_size) {}
  float threshold() const { return m_threshold; }
  size_t window() const { return m_window; }

 private:
  float m_threshold;
  size_t m_window;
};

// helpers/BitVector.h:361-400 and 415-471 (host versions of the per-voxel steps the
// GPU kernels apply)
void performLeftShift(BitVector<BIT_VECTOR_LENGTH>& bit_vector, const uint8_t shift_size);
bool bitMarginCollisionCheck(const BitVector<BIT_VECTOR_LENGTH>& v1, const BitVector<BIT_VECTOR_LENGTH>& v2,
                             BitVector<BIT_VECTOR_LENGTH>* collisions, const uint8_t margin, const uint32_t sv_offset);

class GpuVoxelsMap;
typedef std::shared_ptr<GpuVoxelsMap> GpuVoxelsMapSharedPtr;

namespace voxelmap {
class BitVectorVoxelMap;
class ProbVoxelMap;
class DistanceVoxelMap;
}  // namespace voxelmap
namespace voxellist {
class BitVectorVoxelList;
}

// ---------------------------------------------------------------- GpuVoxelsMap.h
class GpuVoxelsMap {
 public:
  virtual ~GpuVoxelsMap() {}
  virtual void insertPointCloud(const std::vector<Vector3f>& point_cloud, const BitVoxelMeaning voxel_meaning) = 0;
  virtual void insertPointCloud(const PointCloud& pointcloud, const BitVoxelMeaning voxel_meaning)
  {
    insertPointCloud(pointcloud.points(), voxel_meaning);
  }
  // d_points: device pointer
  virtual void insertPointCloud(const Vector3f* d_points, uint32_t size, const BitVoxelMeaning voxel_meaning) = 0;
  virtual void insertCoordinateList(const std::vector<Vector3ui>& coordinates, const BitVoxelMeaning voxel_meaning) = 0;
  virtual void insertMetaPointCloud(const MetaPointCloud& meta_point_cloud, BitVoxelMeaning voxel_meaning) = 0;
  virtual void insertMetaPointCloud(const MetaPointCloud& meta_point_cloud,
                                    const std::vector<BitVoxelMeaning>& voxel_meanings) = 0;
  // GpuVoxelsMap.h:141-153 (BitVoxelMap.hpp:228-341): each sub-cloud gets its meaning
  // (default eBVM_SWEPT_VOLUME_START + i) and collides where its voxels already hold a bit
  // of its mask (default: all but eBVM_COLLISION and its own meaning).  Only BitVector maps
  // support it; the other maps report "not supported" and return true, as the reference.
  virtual bool insertMetaPointCloudWithSelfCollisionCheck(
      const MetaPointCloud* robot_links, const std::vector<BitVoxelMeaning>& voxel_meanings = std::vector<BitVoxelMeaning>(),
      const std::vector<BitVector<BIT_VECTOR_LENGTH> >& collision_masks = std::vector<BitVector<BIT_VECTOR_LENGTH> >(),
      BitVector<BIT_VECTOR_LENGTH>* colliding_meanings = NULL);
  virtual bool merge(const GpuVoxelsMapSharedPtr other, const Vector3f& metric_offset = Vector3f(),
                     const BitVoxelMeaning* new_meaning = NULL);
  virtual bool merge(const GpuVoxelsMapSharedPtr other, const Vector3i& voxel_offset,
                     const BitVoxelMeaning* new_meaning = NULL);
  virtual void clearMap() = 0;
  virtual void clearBitVoxelMeaning(BitVoxelMeaning voxel_meaning);
  virtual std::size_t getMemoryUsage() const = 0;
  virtual Vector3ui getDimensions() const = 0;
  virtual float getVoxelSideLength() const { return m_side; }
  MapType getMapType() const { return m_map_type; }
  template <class T>
  T* as() { return dynamic_cast<T*>(this); }
  template <class T>
  const T* as() const { return dynamic_cast<const T*>(this); }
  // GpuVoxelsMap.cpp:36-48: load a .xyz / .pcd / .binvox file and insert its points
  bool insertPointCloudFromFile(const std::string path, const bool use_model_path, const BitVoxelMeaning voxel_meaning,
                                const bool shift_to_zero = false, const Vector3f& offset_XYZ = Vector3f(),
                                const float scaling = 1.0);
  // map / list dumps (TemplateVoxelMap.hpp:906-1039, TemplateVoxelList.hpp:625-715)
  virtual bool writeToDisk(const std::string path) = 0;
  virtual bool readFromDisk(const std::string path) = 0;
  virtual bool needsRebuild() const { return false; }
  virtual bool rebuild() { return true; }

 protected:
  MapType m_map_type = MT_BITVECTOR_VOXELMAP;
  float m_side = 0.f;
};

namespace voxelmap {

class TemplateVoxelMapBase : public GpuVoxelsMap {
 public:
  TemplateVoxelMapBase(const Vector3ui dim, const float voxel_side_length, const MapType map_type);
  ~TemplateVoxelMapBase() override;
  void insertPointCloud(const std::vector<Vector3f>& point_cloud, const BitVoxelMeaning voxel_meaning) override;
  void insertPointCloud(const Vector3f* d_points, uint32_t size, const BitVoxelMeaning voxel_meaning) override;
  using GpuVoxelsMap::insertPointCloud;
  void insertCoordinateList(const std::vector<Vector3ui>& coordinates, const BitVoxelMeaning voxel_meaning) override;
  void insertMetaPointCloud(const MetaPointCloud& meta_point_cloud, BitVoxelMeaning voxel_meaning) override;
  void insertMetaPointCloud(const MetaPointCloud& meta_point_cloud,
                            const std::vector<BitVoxelMeaning>& voxel_meanings) override;
  void clearMap() override;
  bool writeToDisk(const std::string path) override;
  bool readFromDisk(const std::string path) override;
  std::size_t getMemoryUsage() const override;
  Vector3ui getDimensions() const override { return m_dim; }
  uint32_t getVoxelMapSize() const { return m_dim.x * m_dim.y * m_dim.z; }
  void* getVoidDeviceDataPtr() const;
  gvx_map handle() const { return m_h; }
  bool lastInsertHadPointsOutside() const { return m_outside; }

 protected:
  gvx_map m_h = nullptr;
  Vector3ui m_dim;
  bool m_outside = false;
};

class BitVectorVoxelMap : public TemplateVoxelMapBase {
 public:
  BitVectorVoxelMap(const Vector3ui dim, const float voxel_side_length, const MapType map_type = MT_BITVECTOR_VOXELMAP)
      : TemplateVoxelMapBase(dim, voxel_side_length, map_type) {}
  bool insertMetaPointCloudWithSelfCollisionCheck(
      const MetaPointCloud* robot_links, const std::vector<BitVoxelMeaning>& voxel_meanings = std::vector<BitVoxelMeaning>(),
      const std::vector<BitVector<BIT_VECTOR_LENGTH> >& collision_masks = std::vector<BitVector<BIT_VECTOR_LENGTH> >(),
      BitVector<BIT_VECTOR_LENGTH>* colliding_meanings = NULL) override;
  // BitVoxelMap.hpp:199-224
  std::size_t collideWith(const BitVectorVoxelMap* map, float coll_threshold = 1.0, const Vector3i& offset = Vector3i());
  std::size_t collideWith(const ProbVoxelMap* map, float coll_threshold = 1.0, const Vector3i& offset = Vector3i());
  std::size_t collideWithTypes(const BitVectorVoxelMap* map, BitVectorVoxel& types_in_collision,
                               float coll_threshold = 1.0, const Vector3i& offset = Vector3i());
  std::size_t collideWithTypes(const ProbVoxelMap* map, BitVectorVoxel& types_in_collision,
                               float coll_threshold = 1.0, const Vector3i& offset = Vector3i());
  // BitVoxelMap.hpp:90-195: colliding_meanings |= OR of the colliding voxels' results
  uint32_t collisionCheckBitvector(const BitVectorVoxelMap* other, SVCollider collider,
                                   BitVector<BIT_VECTOR_LENGTH>& colliding_meanings, const uint16_t sv_offset = 0);
  uint32_t collisionCheckBitvector(const ProbVoxelMap* other, SVCollider collider,
                                   BitVector<BIT_VECTOR_LENGTH>& colliding_meanings, const uint16_t sv_offset = 0);
  // BitVoxelMap.hpp:351-362 (shift_size <= 56)
  void shiftLeftSweptVolumeIDs(uint8_t shift_size);
  void clearBit(const uint32_t bit_index);
  void clearBits(BitVector<BIT_VECTOR_LENGTH> bits);
  void clearBitVoxelMeaning(BitVoxelMeaning voxel_meaning) override { clearBit(voxel_meaning); }
  BitVectorVoxel* getDeviceDataPtr() const { return (BitVectorVoxel*)getVoidDeviceDataPtr(); }
};

class ProbVoxelMap : public TemplateVoxelMapBase {
 public:
  ProbVoxelMap(const Vector3ui dim, const float voxel_side_length, const MapType map_type = MT_PROBAB_VOXELMAP)
      : TemplateVoxelMapBase(dim, voxel_side_length, map_type) {}
  // ProbVoxelMap.hpp:98-109
  std::size_t collideWith(const BitVectorVoxelMap* map, float coll_threshold = 1.0, const Vector3i& offset = Vector3i());
  std::size_t collideWith(const ProbVoxelMap* map, float coll_threshold = 1.0, const Vector3i& offset = Vector3i());
  void clearBitVoxelMeaning(BitVoxelMeaning) override { clearMap(); }
  Probability* getDeviceDataPtr() const { return (Probability*)getVoidDeviceDataPtr(); }
  // ProbVoxelMap.hpp:51-78: ray-cast one scan from the sensor voxel (FREE along every
  // ray, then OCCUPIED at every measured voxel); robot_map: device voxels of a
  // BitVectorVoxelMap of the same size (getDeviceDataPtr), used when cut_real_robot
  void insertSensorData(const PointCloud& global_points, const Vector3f& sensor_pose, const bool enable_raycasting,
                        const bool cut_real_robot, const BitVoxelMeaning robot_voxel_meaning,
                        BitVectorVoxel* robot_map);
};

class DistanceVoxelMap : public TemplateVoxelMapBase {
 public:
  DistanceVoxelMap(const Vector3ui dim, const float voxel_side_length, const MapType map_type = MT_DISTANCE_VOXELMAP)
      : TemplateVoxelMapBase(dim, voxel_side_length, map_type) {}
  // DistanceVoxelMap.hpp:109-136
  bool mergeOccupied(const std::shared_ptr<ProbVoxelMap> other, const Vector3ui& voxel_offset = Vector3ui(),
                     float occupancy_threshold = 0.5);
  // DistanceVoxelMap.hpp:284-645; the banding arguments do not change the result
  void parallelBanding3D(uint32_t m1 = 1, uint32_t m2 = 1, uint32_t m3 = 1, uint32_t m1_blocksize = 64,
                         uint32_t m2_blocksize = 64, uint32_t m3_blocksize = 16, bool detailtimer = false);
  // DistanceVoxelMap.hpp:208-253
  void exactDistances3D(std::vector<Vector3f>& points);
  int32_t getSquaredObstacleDistance(const Vector3ui& pos);
  int32_t getSquaredObstacleDistance(uint32_t x, uint32_t y, uint32_t z);
  int32_t getObstacleDistance(const Vector3ui& pos);
  int32_t getObstacleDistance(uint32_t x, uint32_t y, uint32_t z);
  void getSquaredDistancesToHost(std::vector<uint32_t>& indices, std::vector<int32_t>& output);
  void getSquaredDistances(std::vector<int32_t>& output);
  void clearBitVoxelMeaning(BitVoxelMeaning) override {}
  uint32_t* getDeviceDataPtr() const { return (uint32_t*)getVoidDeviceDataPtr(); }
};

}  // namespace voxelmap

namespace voxellist {

class BitVectorVoxelList : public GpuVoxelsMap {
 public:
  BitVectorVoxelList(const Vector3ui ref_map_dim, const float voxel_side_length,
                     const MapType map_type = MT_BITVECTOR_VOXELLIST);
  ~BitVectorVoxelList() override;
  void insertPointCloud(const std::vector<Vector3f>& point_cloud, const BitVoxelMeaning voxel_meaning) override;
  void insertPointCloud(const Vector3f* d_points, uint32_t size, const BitVoxelMeaning voxel_meaning) override;
  using GpuVoxelsMap::insertPointCloud;
  void insertCoordinateList(const std::vector<Vector3ui>& coordinates, const BitVoxelMeaning voxel_meaning) override;
  void insertMetaPointCloud(const MetaPointCloud& meta_point_cloud, BitVoxelMeaning voxel_meaning) override;
  void insertMetaPointCloud(const MetaPointCloud& meta_point_cloud,
                            const std::vector<BitVoxelMeaning>& voxel_meanings) override;
  bool merge(const GpuVoxelsMapSharedPtr other, const Vector3f& metric_offset = Vector3f(),
             const BitVoxelMeaning* new_meaning = NULL) override;
  bool merge(const GpuVoxelsMapSharedPtr other, const Vector3i& voxel_offset,
             const BitVoxelMeaning* new_meaning = NULL) override;
  void clearMap() override;
  bool writeToDisk(const std::string path) override;
  bool readFromDisk(const std::string path) override;
  std::size_t getMemoryUsage() const override;
  // lists report their length in x (TemplateVoxelList.h)
  Vector3ui getDimensions() const override;
  Vector3ui getMetaDimensions() const { return m_ref_dim; }
  // BitVoxelList.hpp:63-258
  std::size_t collideWith(const voxelmap::BitVectorVoxelMap* map, float coll_threshold = 1.0,
                          const Vector3i& offset = Vector3i());
  std::size_t collideWith(const voxelmap::ProbVoxelMap* map, float coll_threshold = 1.0,
                          const Vector3i& offset = Vector3i());
  std::size_t collideWith(const BitVectorVoxelList* other, float coll_threshold = 1.0,
                          const Vector3i& offset = Vector3i());
  std::size_t collideWithTypes(const voxelmap::BitVectorVoxelMap* map, BitVectorVoxel& types_in_collision,
                               float coll_threshold = 1.0, const Vector3i& offset = Vector3i());
  std::size_t collideWithTypes(const voxelmap::ProbVoxelMap* map, BitVectorVoxel& types_in_collision,
                               float coll_threshold = 1.0, const Vector3i& offset = Vector3i());
  std::size_t collideWithTypes(const BitVectorVoxelList* other, BitVectorVoxel& types_in_collision,
                               float coll_threshold = 1.0, const Vector3i& offset = Vector3i());
  std::size_t collideWithTypeMask(const voxelmap::BitVectorVoxelMap* map, const BitVectorVoxel& types_to_check,
                                  float coll_threshold = 1.0, const Vector3i& offset = Vector3i());
  std::size_t collideWithTypeMask(const voxelmap::ProbVoxelMap* map, const BitVectorVoxel& types_to_check,
                                  float coll_threshold = 1.0, const Vector3i& offset = Vector3i());
  // BitVoxelList.hpp:260-292, 295-343, 641-664
  std::size_t collideWithBitcheck(const BitVectorVoxelList* other, const uint8_t margin = 0,
                                  const Vector3i& offset = Vector3i());
  std::size_t collideCountingPerMeaning(const GpuVoxelsMapSharedPtr other, std::vector<std::size_t>& collisions_per_meaning,
                                        const Vector3i& offset = Vector3i());
  void shiftLeftSweptVolumeIDs(uint8_t shift_size);
  void copyCoordsToHost(std::vector<Vector3ui>& host_vec);
  void copyVoxelsToHost(std::vector<uint32_t>& ids, std::vector<Vector3ui>& coords, std::vector<BitVectorVoxel>& voxels);
  uint32_t* getDeviceIdPtr() const;
  Vector3ui* getDeviceCoordPtr() const;
  BitVectorVoxel* getDeviceDataPtr() const;
  gvx_list handle() const { return m_h; }

 private:
  std::size_t collideMap(gvx_map map, float thr, const Vector3i& offset, int mode, const uint8_t* typemask,
                         BitVectorVoxel* types);
  gvx_list m_h = nullptr;
  Vector3ui m_ref_dim;
};

}  // namespace voxellist

namespace robot {

typedef std::map<std::string, float> JointValueMap;
typedef enum { REVOLUTE = 0, PRISMATIC } DHJointType;

// robot/dh_robot/KinematicLink.h
class DHParameters {
 public:
  float d, theta, a, alpha, value;
  DHJointType joint_type;
  DHParameters() : d(0), theta(0), a(0), alpha(0), value(0), joint_type(REVOLUTE) {}
  DHParameters(float _d, float _theta, float _a, float _alpha, float _value, DHJointType _joint_type)
      : d(_d), theta(_theta), a(_a), alpha(_alpha), value(_value), joint_type(_joint_type) {}
  // DH -> homogeneous transform (KinematicLink.cu:24-79), evaluated on the device's
  // deterministic sine/cosine so host and GPU agree bit for bit
  void convertDHtoM(Matrix4f& m) const;
};

// robot/dh_robot/KinematicChain.h: link cloud i is posed by T_0 ... T_{i-1}
class KinematicChain {
 public:
  KinematicChain(const std::vector<std::string>& linknames, const std::vector<DHParameters>& dh_params,
                 const MetaPointCloud& pointclouds, const Matrix4f& base = Matrix4f::createIdentity());
  // KinematicChain.cu:35-43: link clouds read from files (named after the files)
  KinematicChain(const std::vector<std::string>& linknames, const std::vector<DHParameters>& dh_params,
                 const std::vector<std::string>& paths_to_pointclouds, const bool use_model_path);
  void setConfiguration(const JointValueMap& jointmap);
  void getConfiguration(JointValueMap& jointmap) const;
  // the posed link clouds of the current configuration
  const MetaPointCloud* getTransformedClouds();
  void setBaseTransformation(const Matrix4f& base) { m_base = base; }
  // fused FK + transform + insert (one GPU launch for all links)
  bool insertIntoMap(GpuVoxelsMap* map, BitVoxelMeaning meaning);
  bool insertIntoMap(GpuVoxelsMap* map, const std::vector<BitVoxelMeaning>& meanings);
  // setConfiguration + insertMetaPointCloudWithSelfCollisionCheck of the posed links, fused
  bool insertIntoMapSelfCollAware(GpuVoxelsMap* map, const std::vector<BitVoxelMeaning>& meanings,
                                  const std::vector<BitVector<BIT_VECTOR_LENGTH> >& masks,
                                  BitVector<BIT_VECTOR_LENGTH>* colliding_meanings);
  uint32_t getNumberOfLinks() const { return (uint32_t)m_names.size(); }

 private:
  void packed(std::vector<float>& dh, std::vector<float>& q, float base[12]) const;
  std::vector<std::string> m_names;
  std::vector<DHParameters> m_dh;
  MetaPointCloud m_clouds;
  MetaPointCloud m_transformed;
  Matrix4f m_base;
};
typedef std::shared_ptr<KinematicChain> RobotInterfaceSharedPtr;

}  // namespace robot

// ---------------------------------------------------------------- GpuVoxels.h
class GpuVoxels;
typedef std::shared_ptr<GpuVoxels> GpuVoxelsSharedPtr;

class GpuVoxels {
 public:
  static GpuVoxelsSharedPtr getInstance();
  ~GpuVoxels();
  void initialize(const uint32_t dim_x, const uint32_t dim_y, const uint32_t dim_z, const float voxel_side_length);
  GpuVoxelsMapSharedPtr addMap(const MapType map_type, const std::string& map_name);
  bool delMap(const std::string& map_name);
  bool clearMap(const std::string& map_name);
  bool clearMap(const std::string& map_name, BitVoxelMeaning voxel_meaning);
  GpuVoxelsMapSharedPtr getMap(const std::string& map_name);
  bool visualizeMap(const std::string&, const bool = true) { return false; }
  bool addRobot(const std::string& robot_name, const std::vector<std::string>& link_names,
                const std::vector<robot::DHParameters>& dh_params, const MetaPointCloud& pointclouds);
  bool addRobot(const std::string& robot_name, const std::vector<std::string>& link_names,
                const std::vector<robot::DHParameters>& dh_params,
                const std::vector<std::string>& paths_to_pointclouds, const bool use_model_path);
  bool insertPointCloudFromFile(const std::string map_name, const std::string path, const bool use_model_path,
                                const BitVoxelMeaning voxel_meaning, const bool shift_to_zero = false,
                                const Vector3f& offset_XYZ = Vector3f(), const float scaling = 1.0);
  robot::RobotInterfaceSharedPtr getRobot(const std::string& rob_name);
  bool setRobotConfiguration(std::string robot_name, const robot::JointValueMap& jointmap);
  bool getRobotConfiguration(const std::string& robot_name, robot::JointValueMap& jointmap);
  bool insertPointCloudIntoMap(const PointCloud& cloud, std::string map_name, const BitVoxelMeaning voxel_meaning);
  bool insertPointCloudIntoMap(const std::vector<Vector3f>& cloud, std::string map_name,
                               const BitVoxelMeaning voxel_meaning);
  bool insertMetaPointCloudIntoMap(const MetaPointCloud& meta_point_cloud, std::string map_name,
                                   const std::vector<BitVoxelMeaning>& voxel_meanings);
  bool insertMetaPointCloudIntoMap(const MetaPointCloud& meta_point_cloud, std::string map_name,
                                   const BitVoxelMeaning voxel_meaning);
  bool insertRobotIntoMap(std::string robot_name, std::string map_name, const BitVoxelMeaning voxel_meaning);
  // GpuVoxels.cpp:519-541: true when a link collided with previously inserted data
  bool insertRobotIntoMapSelfCollAware(
      std::string robot_name, std::string map_name,
      const std::vector<BitVoxelMeaning>& voxel_meanings = std::vector<BitVoxelMeaning>(),
      const std::vector<BitVector<BIT_VECTOR_LENGTH> >& collision_masks = std::vector<BitVector<BIT_VECTOR_LENGTH> >(),
      BitVector<BIT_VECTOR_LENGTH>* colliding_meanings = NULL);
  bool insertBoxIntoMap(const Vector3f& corner_min, const Vector3f& corner_max, std::string map_name,
                        const BitVoxelMeaning voxel_meaning, uint16_t points_per_voxel = 1);
  void getDimensions(uint32_t& dim_x, uint32_t& dim_y, uint32_t& dim_z);
  void getDimensions(Vector3ui& dim);
  void getVoxelSideLength(float& voxel_side_length);

 private:
  GpuVoxels();
  std::map<std::string, GpuVoxelsMapSharedPtr> m_managed_maps;
  std::map<std::string, robot::RobotInterfaceSharedPtr> m_managed_robots;
  Vector3ui m_dim;
  float m_voxel_side_length = 0.f;
  bool m_initialized = false;
};

}  // namespace gpu_voxels

#endif
