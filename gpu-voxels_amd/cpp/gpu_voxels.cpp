// gpu_voxels.cpp -- the C++ drop-in API (include/gpu_voxels/GpuVoxels.h) over the C ABI.
#include "../../include/gpu_voxels/GpuVoxels.h"

#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>

namespace gpu_voxels {

namespace {

void warn(const char* where, int code)
{
  std::fprintf(stderr, "gpu_voxels: %s failed (%d): %s\n", where, code, gvx_last_error());
}

bool ok(int code, const char* where)
{
  if (code != GVX_OK) {
    warn(where, code);
    return false;
  }
  return true;
}

void ensure_init()
{
  static std::once_flag once;
  std::call_once(once, []() {
    int s = gvx_init(0);
    if (s != GVX_OK) {
      warn("gvx_init", s);
      throw std::runtime_error(std::string("GPU init failed: ") + gvx_last_error());
    }
  });
}

std::vector<uint32_t> to_u32(const std::vector<BitVoxelMeaning>& m)
{
  std::vector<uint32_t> r(m.size());
  for (std::size_t i = 0; i < m.size(); ++i) r[i] = (uint32_t)m[i];
  return r;
}

const int32_t* off3(const Vector3i& o, int32_t buf[3])
{
  buf[0] = o.x; buf[1] = o.y; buf[2] = o.z;
  return buf;
}

}  // namespace

// ---------------------------------------------------------------- scalar helpers
Probability ProbabilisticVoxel::floatToProbability(const float val)
{
  float tmp = val < 1.0f ? val : 1.0f;
  tmp = tmp > 0.0f ? tmp : 0.0f;
  volatile float prod = tmp * (float(MAX_PROBABILITY) - float(MIN_PROBABILITY));
  tmp = prod + float(MIN_PROBABILITY);
  return Probability(tmp);
}

float ProbabilisticVoxel::probabilityToFloat(const Probability val)
{
  return (float(val) - float(MIN_PROBABILITY)) / (float(MAX_PROBABILITY) - float(MIN_PROBABILITY));
}

MetaPointCloud::MetaPointCloud(const std::vector<std::vector<Vector3f> >& clouds)
{
  for (const auto& c : clouds) addCloud(c);
}

MetaPointCloud::MetaPointCloud(const std::vector<std::vector<Vector3f> >& clouds, const std::vector<std::string>& names)
{
  for (std::size_t i = 0; i < clouds.size(); ++i) addCloud(clouds[i], i < names.size() ? names[i] : "");
}

MetaPointCloud::MetaPointCloud(const std::vector<std::string>& files, bool use_model_path)
    : MetaPointCloud(files, files, use_model_path)
{
}

MetaPointCloud::MetaPointCloud(const std::vector<std::string>& files, const std::vector<std::string>& names,
                               bool use_model_path)
{
  std::vector<std::vector<Vector3f> > clouds(files.size());
  for (std::size_t i = 0; i < files.size(); ++i)
    if (!file_handling::PointcloudFileHandler::Instance()->loadPointCloud(files[i], use_model_path, clouds[i]))
      throw std::runtime_error("MetaPointCloud: could not read " + files[i]);
  for (std::size_t i = 0; i < files.size(); ++i)
    addCloud(clouds[i], names.size() == files.size() ? names[i] : files[i]);
}

void MetaPointCloud::addCloud(const std::vector<Vector3f>& cloud, const std::string& name)
{
  m_points.insert(m_points.end(), cloud.begin(), cloud.end());
  m_sizes.push_back((uint32_t)cloud.size());
  m_names.push_back(name);
}

void MetaPointCloud::updatePointCloud(uint16_t cloud, const std::vector<Vector3f>& pts)
{
  std::size_t off = 0;
  for (uint16_t c = 0; c < cloud; ++c) off += m_sizes[c];
  m_points.erase(m_points.begin() + off, m_points.begin() + off + m_sizes[cloud]);
  m_points.insert(m_points.begin() + off, pts.begin(), pts.end());
  m_sizes[cloud] = (uint32_t)pts.size();
}

bool MetaPointCloud::hasCloud(const std::string& name) const { return getCloudNumber(name) >= 0; }

int16_t MetaPointCloud::getCloudNumber(const std::string& name) const
{
  for (std::size_t i = 0; i < m_names.size(); ++i)
    if (m_names[i] == name) return (int16_t)i;
  return -1;
}

// ---------------------------------------------------------------- bit vector steps
void performLeftShift(BitVector<BIT_VECTOR_LENGTH>& v, const uint8_t shift)
{
  uint64_t buffer = 0;
  for (uint32_t k = 0; k < 8; ++k) buffer += (uint64_t)v.getByte(k * 8) << (k * 8);
  for (uint32_t k = 0; k < BIT_VECTOR_LENGTH / 8; ++k) {
    uint8_t b = (uint8_t)(buffer >> shift);
    if (k == 0) b &= 0xF0;  // bits 0..3 are not swept-volume meanings
    v.setByte(k * 8, b);
    buffer >>= 8;
    if (k + 8 < BIT_VECTOR_LENGTH / 8) buffer += (uint64_t)v.getByte((k + 8) * 8) << 56;
  }
}

bool bitMarginCollisionCheck(const BitVector<BIT_VECTOR_LENGTH>& v1, const BitVector<BIT_VECTOR_LENGTH>& v2,
                             BitVector<BIT_VECTOR_LENGTH>* collisions, const uint8_t margin, const uint32_t sv_offset)
{
  uint64_t buffer = 0;
  for (uint32_t k = 0; k < 4; ++k) buffer += (uint64_t)(k == 0 ? (v2.getByte(0) & 0xF0) : v2.getByte(k * 8)) << (32 + 8 * k);
  const uint32_t byte_off = sv_offset % 8, bit_off = sv_offset / 8;
  for (uint32_t i = eBVM_SWEPT_VOLUME_START + byte_off; i < eBVM_SWEPT_VOLUME_END; i += 8) {
    uint8_t byte = 0;
    uint64_t b1 = (uint64_t)v1.getByte(i) << (32 - margin + bit_off);
    for (uint32_t s = 0; s <= 2u * margin; ++s) {
      byte |= (uint8_t)((b1 & buffer) >> (32 - margin + s));
      b1 <<= 1;
    }
    collisions->setByte(i, byte);
    buffer >>= 8;
    if (i < BIT_VECTOR_LENGTH - 32) buffer += (uint64_t)v2.getByte(i + 32) << 56;
  }
  return !collisions->isZero();
}

namespace file_handling {

PointcloudFileHandler* PointcloudFileHandler::Instance()
{
  static PointcloudFileHandler instance;
  return &instance;
}

bool PointcloudFileHandler::loadPointCloud(const std::string path, const bool use_model_path,
                                           std::vector<Vector3f>& points, const bool shift_to_zero,
                                           const Vector3f& offset_XYZ, const float scaling)
{
  float* xyz = nullptr;
  uint64_t n = 0;
  const float off[3] = {offset_XYZ.x, offset_XYZ.y, offset_XYZ.z};
  if (!ok(gvx_load_point_cloud(path.c_str(), use_model_path ? 1 : 0, shift_to_zero ? 1 : 0, off, scaling, &xyz, &n),
          "loadPointCloud"))
    return false;
  const std::size_t base = points.size();
  points.resize(base + n);
  for (uint64_t i = 0; i < n; ++i) points[base + i] = Vector3f(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
  gvx_free(xyz);
  return true;
}

// PointcloudFileHandler.cpp:144-171: subtract the centre of the bounding box
void PointcloudFileHandler::centerPointCloud(std::vector<Vector3f>& points)
{
  if (points.empty()) return;
  Vector3f lo = points[0], hi = points[0];
  for (const Vector3f& p : points) {
    lo.x = std::min(lo.x, p.x); lo.y = std::min(lo.y, p.y); lo.z = std::min(lo.z, p.z);
    hi.x = std::max(hi.x, p.x); hi.y = std::max(hi.y, p.y); hi.z = std::max(hi.z, p.z);
  }
  const Vector3f c((float)((lo.x + hi.x) / 2.0), (float)((lo.y + hi.y) / 2.0), (float)((lo.z + hi.z) / 2.0));
  for (Vector3f& p : points) { p.x -= c.x; p.y -= c.y; p.z -= c.z; }
}

// PointcloudFileHandler.cpp:177-194: move the minimum corner to zero
void PointcloudFileHandler::shiftPointCloudToZero(std::vector<Vector3f>& points)
{
  if (points.empty()) return;
  Vector3f lo = points[0];
  for (const Vector3f& p : points) { lo.x = std::min(lo.x, p.x); lo.y = std::min(lo.y, p.y); lo.z = std::min(lo.z, p.z); }
  for (Vector3f& p : points) { p.x -= lo.x; p.y -= lo.y; p.z -= lo.z; }
}

}  // namespace file_handling

bool GpuVoxelsMap::insertPointCloudFromFile(const std::string path, const bool use_model_path,
                                            const BitVoxelMeaning voxel_meaning, const bool shift_to_zero,
                                            const Vector3f& offset_XYZ, const float scaling)
{
  std::vector<Vector3f> points;
  if (!file_handling::PointcloudFileHandler::Instance()->loadPointCloud(path, use_model_path, points, shift_to_zero,
                                                                        offset_XYZ, scaling))
    return false;
  insertPointCloud(points, voxel_meaning);
  return true;
}

namespace geometry_generation {

std::vector<Vector3f> createBoxOfPoints(Vector3f min, Vector3f max, float delta)
{
  std::vector<Vector3f> box;
  for (float x = min.x; x <= max.x; x += delta)
    for (float y = min.y; y <= max.y; y += delta)
      for (float z = min.z; z <= max.z; z += delta) box.push_back(Vector3f(x, y, z));
  return box;
}

std::vector<Vector3ui> createBoxOfPoints(Vector3f min, Vector3f max, float delta, float side)
{
  Vector3f mc(min.x >= 0.0f ? min.x : 0.0f, min.y >= 0.0f ? min.y : 0.0f, min.z >= 0.0f ? min.z : 0.0f);
  Vector3ui lo((uint32_t)std::floor(mc.x / side), (uint32_t)std::floor(mc.y / side), (uint32_t)std::floor(mc.z / side));
  Vector3ui hi((uint32_t)std::ceil(max.x / side), (uint32_t)std::ceil(max.y / side), (uint32_t)std::ceil(max.z / side));
  uint32_t d = (uint32_t)std::round(delta / side);
  if (d == 0) d = 1;
  std::vector<Vector3ui> out;
  for (float x = (float)lo.x; x <= hi.x; x += d)
    for (float y = (float)lo.y; y <= hi.y; y += d)
      for (float z = (float)lo.z; z <= hi.z; z += d) out.push_back(Vector3ui((uint32_t)x, (uint32_t)y, (uint32_t)z));
  return out;
}

}  // namespace geometry_generation

// ---------------------------------------------------------------- GpuVoxelsMap defaults
bool GpuVoxelsMap::merge(const GpuVoxelsMapSharedPtr, const Vector3f&, const BitVoxelMeaning*)
{
  std::fprintf(stderr, "gpu_voxels: %s\n", GPU_VOXELS_MAP_OPERATION_NOT_SUPPORTED.c_str());
  return false;
}

bool GpuVoxelsMap::merge(const GpuVoxelsMapSharedPtr, const Vector3i&, const BitVoxelMeaning*)
{
  std::fprintf(stderr, "gpu_voxels: %s\n", GPU_VOXELS_MAP_OPERATION_NOT_SUPPORTED.c_str());
  return false;
}

void GpuVoxelsMap::clearBitVoxelMeaning(BitVoxelMeaning) {}

bool GpuVoxelsMap::insertMetaPointCloudWithSelfCollisionCheck(const MetaPointCloud*,
                                                              const std::vector<BitVoxelMeaning>&,
                                                              const std::vector<BitVector<BIT_VECTOR_LENGTH> >&,
                                                              BitVector<BIT_VECTOR_LENGTH>*)
{
  std::fprintf(stderr, "gpu_voxels: insertMetaPointCloudWithSelfCollisionCheck is not supported by this map type\n");
  return true;
}

// meanings / masks of the self-collision-checking inserts: empty = the defaults, else one per sub-cloud
static bool selfcoll_args(size_t n, const std::vector<BitVoxelMeaning>& ms,
                          const std::vector<BitVector<BIT_VECTOR_LENGTH> >& masks, std::vector<uint32_t>& mn,
                          std::vector<uint8_t>& mk)
{
  if ((!ms.empty() && ms.size() != n) || (!masks.empty() && masks.size() != n)) {
    std::fprintf(stderr, "gpu_voxels: meanings / collision masks need one entry per sub-cloud\n");
    return false;
  }
  mn = to_u32(ms);
  mk.clear();
  for (const auto& b : masks) mk.insert(mk.end(), b.data(), b.data() + BIT_VECTOR_LENGTH / 8);
  return true;
}

// ---------------------------------------------------------------- dense maps
namespace voxelmap {

static int gvx_type(MapType t)
{
  switch (t) {
    case MT_BITVECTOR_VOXELMAP: return GVX_MT_BITVECTOR_VOXELMAP;
    case MT_PROBAB_VOXELMAP: return GVX_MT_PROBAB_VOXELMAP;
    case MT_DISTANCE_VOXELMAP: return GVX_MT_DISTANCE_VOXELMAP;
    default: throw GPU_VOXELS_MAP_TYPE_NOT_IMPLEMENTED;
  }
}

TemplateVoxelMapBase::TemplateVoxelMapBase(const Vector3ui dim, const float side, const MapType map_type) : m_dim(dim)
{
  ensure_init();
  m_map_type = map_type;
  m_side = side;
  int s = gvx_map_create(gvx_type(map_type), dim.x, dim.y, dim.z, side, &m_h);
  if (s != GVX_OK) {
    warn("gvx_map_create", s);
    throw std::runtime_error(std::string("voxel map allocation failed: ") + gvx_last_error());
  }
}

TemplateVoxelMapBase::~TemplateVoxelMapBase() { gvx_map_destroy(m_h); }

void TemplateVoxelMapBase::insertPointCloud(const std::vector<Vector3f>& pc, const BitVoxelMeaning m)
{
  int out = 0;
  if (ok(gvx_map_insert_points(m_h, pc.empty() ? nullptr : &pc[0].x, pc.size(), (uint32_t)m, GVX_MEM_HOST, &out),
         "insertPointCloud"))
    m_outside = out != 0;
  if (m_outside) std::fprintf(stderr, "gpu_voxels: You tried to insert points that lie outside the map dimensions!\n");
}

void TemplateVoxelMapBase::insertPointCloud(const Vector3f* d_points, uint32_t size, const BitVoxelMeaning m)
{
  int out = 0;
  if (ok(gvx_map_insert_points(m_h, &d_points->x, size, (uint32_t)m, GVX_MEM_DEVICE, &out), "insertPointCloud"))
    m_outside = out != 0;
}

void TemplateVoxelMapBase::insertCoordinateList(const std::vector<Vector3ui>& c, const BitVoxelMeaning m)
{
  int out = 0;
  if (ok(gvx_map_insert_coords(m_h, c.empty() ? nullptr : &c[0].x, c.size(), (uint32_t)m, GVX_MEM_HOST, &out),
         "insertCoordinateList"))
    m_outside = out != 0;
}

void TemplateVoxelMapBase::insertMetaPointCloud(const MetaPointCloud& mpc, BitVoxelMeaning m)
{
  int out = 0;
  std::vector<uint32_t> sizes = mpc.getPointcloudSizes();
  const auto& p = mpc.points();
  if (ok(gvx_map_insert_meta(m_h, p.empty() ? nullptr : &p[0].x, sizes.data(), (uint32_t)sizes.size(), nullptr,
                             (uint32_t)m, GVX_MEM_HOST, &out),
         "insertMetaPointCloud"))
    m_outside = out != 0;
}

void TemplateVoxelMapBase::insertMetaPointCloud(const MetaPointCloud& mpc, const std::vector<BitVoxelMeaning>& ms)
{
  if (ms.size() != mpc.getNumberOfPointclouds()) {
    std::fprintf(stderr, "gpu_voxels: number of meanings differs from number of sub-clouds, not inserting\n");
    return;
  }
  int out = 0;
  std::vector<uint32_t> sizes = mpc.getPointcloudSizes(), mn = to_u32(ms);
  const auto& p = mpc.points();
  if (ok(gvx_map_insert_meta(m_h, p.empty() ? nullptr : &p[0].x, sizes.data(), (uint32_t)sizes.size(), mn.data(), 0,
                             GVX_MEM_HOST, &out),
         "insertMetaPointCloud"))
    m_outside = out != 0;
}

void TemplateVoxelMapBase::clearMap() { ok(gvx_map_clear(m_h), "clearMap"); }

bool BitVectorVoxelMap::insertMetaPointCloudWithSelfCollisionCheck(
    const MetaPointCloud* mpc, const std::vector<BitVoxelMeaning>& ms,
    const std::vector<BitVector<BIT_VECTOR_LENGTH> >& masks, BitVector<BIT_VECTOR_LENGTH>* colliding)
{
  if (!mpc) return false;
  std::vector<uint32_t> sizes = mpc->getPointcloudSizes(), mn;
  std::vector<uint8_t> mk;
  if (!selfcoll_args(sizes.size(), ms, masks, mn, mk)) return false;
  const auto& p = mpc->points();
  BitVector<BIT_VECTOR_LENGTH> col;
  int self = 0, out = 0;
  if (!ok(gvx_map_insert_meta_selfcoll(m_h, p.empty() ? nullptr : &p[0].x, sizes.data(), (uint32_t)sizes.size(),
                                       mn.empty() ? nullptr : mn.data(), mk.empty() ? nullptr : mk.data(),
                                       GVX_MEM_HOST, col.data(), &self, &out),
          "insertMetaPointCloudWithSelfCollisionCheck"))
    return false;
  m_outside = out != 0;
  if (colliding) *colliding = col;
  return self != 0;
}

void ProbVoxelMap::insertSensorData(const PointCloud& global_points, const Vector3f& sensor_pose,
                                    const bool enable_raycasting, const bool cut_real_robot,
                                    const BitVoxelMeaning robot_voxel_meaning, BitVectorVoxel* robot_map)
{
  const std::vector<Vector3f>& p = global_points.points();
  const float pose[3] = {sensor_pose.x, sensor_pose.y, sensor_pose.z};
  ok(gvx_map_insert_sensor(m_h, p.empty() ? nullptr : &p[0].x, p.size(), pose, enable_raycasting ? 1 : 0,
                           cut_real_robot ? (const void*)robot_map : nullptr, (uint32_t)robot_voxel_meaning,
                           GVX_MEM_HOST),
     "insertSensorData");
}

bool TemplateVoxelMapBase::writeToDisk(const std::string path) { return ok(gvx_map_write_file(m_h, path.c_str()), "writeToDisk"); }

bool TemplateVoxelMapBase::readFromDisk(const std::string path)
{
  int side_matches = 1;
  if (!ok(gvx_map_read_file(m_h, path.c_str(), &side_matches), "readFromDisk")) return false;
  if (!side_matches)
    std::fprintf(stderr, "gpu_voxels: readFromDisk: the voxel side length in %s differs from the map's; continuing\n",
                 path.c_str());
  return true;
}

std::size_t TemplateVoxelMapBase::getMemoryUsage() const
{
  uint64_t b = 0;
  gvx_map_info(m_h, nullptr, nullptr, nullptr, &b);
  return (std::size_t)b;
}

void* TemplateVoxelMapBase::getVoidDeviceDataPtr() const
{
  void* p = nullptr;
  gvx_map_device_ptr(m_h, &p);
  return p;
}

static std::size_t collide(gvx_map a, gvx_map b, float thr, const Vector3i& off)
{
  int32_t o[3];
  uint64_t c = 0;
  if (!ok(gvx_map_collide(a, b, thr, off3(off, o), &c), "collideWith")) return SSIZE_MAX;
  return (std::size_t)c;
}

static std::size_t collide_types(gvx_map a, gvx_map b, float thr, BitVectorVoxel& types)
{
  uint64_t c = 0;
  uint8_t mask[32];
  if (!ok(gvx_map_collide_types(a, b, thr, &c, mask), "collideWithTypes")) return SSIZE_MAX;
  types.bitVector() |= [&]() {
    BitVector<BIT_VECTOR_LENGTH> bv;
    std::memcpy(bv.data(), mask, 32);
    return bv;
  }();
  return (std::size_t)c;
}

std::size_t BitVectorVoxelMap::collideWith(const BitVectorVoxelMap* m, float thr, const Vector3i& off)
{
  return collide(m_h, m->handle(), thr, off);
}
std::size_t BitVectorVoxelMap::collideWith(const ProbVoxelMap* m, float thr, const Vector3i& off)
{
  return collide(m_h, m->handle(), thr, off);
}
// the reference ignores the offset of collideWithTypes (BitVoxelMap.hpp:213-224)
std::size_t BitVectorVoxelMap::collideWithTypes(const BitVectorVoxelMap* m, BitVectorVoxel& t, float thr, const Vector3i&)
{
  return collide_types(m_h, m->handle(), thr, t);
}
std::size_t BitVectorVoxelMap::collideWithTypes(const ProbVoxelMap* m, BitVectorVoxel& t, float thr, const Vector3i&)
{
  return collide_types(m_h, m->handle(), thr, t);
}

uint32_t BitVectorVoxelMap::collisionCheckBitvector(const BitVectorVoxelMap* other, SVCollider collider,
                                                    BitVector<BIT_VECTOR_LENGTH>& meanings, const uint16_t sv_offset)
{
  uint64_t c = 0;
  uint8_t mask[32];
  if (!ok(gvx_map_collide_sv(m_h, other->handle(), collider.threshold(), (uint32_t)collider.window(), sv_offset, &c,
                             mask),
          "collisionCheckBitvector"))
    return 0;
  BitVector<BIT_VECTOR_LENGTH> bv;
  std::memcpy(bv.data(), mask, 32);
  meanings |= bv;
  return (uint32_t)c;
}

uint32_t BitVectorVoxelMap::collisionCheckBitvector(const ProbVoxelMap* other, SVCollider collider,
                                                    BitVector<BIT_VECTOR_LENGTH>& meanings, const uint16_t sv_offset)
{
  uint64_t c = 0;
  uint8_t mask[32];
  if (!ok(gvx_map_collide_sv(m_h, other->handle(), collider.threshold(), (uint32_t)collider.window(), sv_offset, &c,
                             mask),
          "collisionCheckBitvector"))
    return 0;
  BitVector<BIT_VECTOR_LENGTH> bv;
  std::memcpy(bv.data(), mask, 32);
  meanings |= bv;
  return (uint32_t)c;
}

void BitVectorVoxelMap::shiftLeftSweptVolumeIDs(uint8_t shift_size)
{
  ok(gvx_map_shift_sv(m_h, shift_size), "shiftLeftSweptVolumeIDs");
}

void BitVectorVoxelMap::clearBit(const uint32_t bit_index)
{
  BitVector<BIT_VECTOR_LENGTH> b;
  b.setBit(bit_index);
  clearBits(b);
}

void BitVectorVoxelMap::clearBits(BitVector<BIT_VECTOR_LENGTH> bits) { ok(gvx_map_clear_bits(m_h, bits.data()), "clearBits"); }

std::size_t ProbVoxelMap::collideWith(const BitVectorVoxelMap* m, float thr, const Vector3i& off)
{
  return collide(m_h, m->handle(), thr, off);
}
std::size_t ProbVoxelMap::collideWith(const ProbVoxelMap* m, float thr, const Vector3i& off)
{
  return collide(m_h, m->handle(), thr, off);
}

bool DistanceVoxelMap::mergeOccupied(const std::shared_ptr<ProbVoxelMap> other, const Vector3ui& off, float thr)
{
  uint32_t o[3] = {off.x, off.y, off.z};
  return ok(gvx_dist_merge_occupied(m_h, other->handle(), o, thr), "mergeOccupied");
}

void DistanceVoxelMap::parallelBanding3D(uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, bool)
{
  ok(gvx_dist_pba(m_h), "parallelBanding3D");
}

void DistanceVoxelMap::exactDistances3D(std::vector<Vector3f>& points)
{
  ok(gvx_dist_exact(m_h, points.empty() ? nullptr : &points[0].x, points.size(), GVX_MEM_HOST), "exactDistances3D");
}

int32_t DistanceVoxelMap::getSquaredObstacleDistance(const Vector3ui& p) { return getSquaredObstacleDistance(p.x, p.y, p.z); }

int32_t DistanceVoxelMap::getSquaredObstacleDistance(uint32_t x, uint32_t y, uint32_t z)
{
  uint32_t idx = (z * m_dim.y + y) * m_dim.x + x;
  int32_t out = 0;
  if (!ok(gvx_dist_squared(m_h, &idx, 1, &out, GVX_MEM_HOST), "getSquaredObstacleDistance")) return MAX_OBSTACLE_DISTANCE;
  return out;
}

// int-truncated square root, as DistanceVoxelMap.hpp:702-708
int32_t DistanceVoxelMap::getObstacleDistance(const Vector3ui& p) { return (int32_t)std::sqrt((double)getSquaredObstacleDistance(p)); }
int32_t DistanceVoxelMap::getObstacleDistance(uint32_t x, uint32_t y, uint32_t z)
{
  return (int32_t)std::sqrt((double)getSquaredObstacleDistance(x, y, z));
}

void DistanceVoxelMap::getSquaredDistancesToHost(std::vector<uint32_t>& indices, std::vector<int32_t>& output)
{
  output.resize(indices.size());
  if (indices.empty()) return;
  ok(gvx_dist_squared(m_h, indices.data(), indices.size(), output.data(), GVX_MEM_HOST), "getSquaredDistancesToHost");
}

void DistanceVoxelMap::getSquaredDistances(std::vector<int32_t>& output)
{
  output.resize(getVoxelMapSize());
  ok(gvx_dist_squared(m_h, nullptr, 0, output.data(), GVX_MEM_HOST), "getSquaredDistances");
}

}  // namespace voxelmap

// ---------------------------------------------------------------- voxel list
namespace voxellist {

BitVectorVoxelList::BitVectorVoxelList(const Vector3ui dim, const float side, const MapType map_type) : m_ref_dim(dim)
{
  ensure_init();
  if (map_type != MT_BITVECTOR_VOXELLIST) throw GPU_VOXELS_MAP_TYPE_NOT_IMPLEMENTED;
  m_map_type = map_type;
  m_side = side;
  int s = gvx_list_create(dim.x, dim.y, dim.z, side, &m_h);
  if (s != GVX_OK) throw std::runtime_error(std::string("voxel list creation failed: ") + gvx_last_error());
}

BitVectorVoxelList::~BitVectorVoxelList() { gvx_list_destroy(m_h); }

void BitVectorVoxelList::insertPointCloud(const std::vector<Vector3f>& pc, const BitVoxelMeaning m)
{
  ok(gvx_list_insert_points(m_h, pc.empty() ? nullptr : &pc[0].x, pc.size(), nullptr, 0, nullptr, (uint32_t)m,
                            GVX_MEM_HOST),
     "insertPointCloud");
}

void BitVectorVoxelList::insertPointCloud(const Vector3f* d_points, uint32_t size, const BitVoxelMeaning m)
{
  ok(gvx_list_insert_points(m_h, &d_points->x, size, nullptr, 0, nullptr, (uint32_t)m, GVX_MEM_DEVICE),
     "insertPointCloud");
}

void BitVectorVoxelList::insertCoordinateList(const std::vector<Vector3ui>& c, const BitVoxelMeaning m)
{
  ok(gvx_list_insert_coords(m_h, c.empty() ? nullptr : &c[0].x, c.size(), (uint32_t)m, GVX_MEM_HOST),
     "insertCoordinateList");
}

void BitVectorVoxelList::insertMetaPointCloud(const MetaPointCloud& mpc, BitVoxelMeaning m)
{
  const auto& p = mpc.points();
  ok(gvx_list_insert_points(m_h, p.empty() ? nullptr : &p[0].x, p.size(), nullptr, 0, nullptr, (uint32_t)m,
                            GVX_MEM_HOST),
     "insertMetaPointCloud");
}

void BitVectorVoxelList::insertMetaPointCloud(const MetaPointCloud& mpc, const std::vector<BitVoxelMeaning>& ms)
{
  if (ms.size() != mpc.getNumberOfPointclouds()) {
    std::fprintf(stderr, "gpu_voxels: Number of VoxelMeanings differs from number of Sub-Pointclouds! Not inserting MetaPointCloud!\n");
    return;
  }
  const auto& p = mpc.points();
  std::vector<uint32_t> sizes = mpc.getPointcloudSizes(), mn = to_u32(ms);
  ok(gvx_list_insert_points(m_h, p.empty() ? nullptr : &p[0].x, p.size(), sizes.data(), (uint32_t)sizes.size(),
                            mn.data(), 0, GVX_MEM_HOST),
     "insertMetaPointCloud");
}

bool BitVectorVoxelList::merge(const GpuVoxelsMapSharedPtr other, const Vector3f& metric_offset,
                               const BitVoxelMeaning* new_meaning)
{
  // mapToVoxelsSigned (VoxelMapOperations.h:134-144)
  Vector3i o((int32_t)std::floor(metric_offset.x / m_side), (int32_t)std::floor(metric_offset.y / m_side),
             (int32_t)std::floor(metric_offset.z / m_side));
  return merge(other, o, new_meaning);
}

bool BitVectorVoxelList::merge(const GpuVoxelsMapSharedPtr other, const Vector3i& off, const BitVoxelMeaning* nm)
{
  int32_t o[3];
  int meaning = nm ? (int)*nm : -1;
  if (auto* l = dynamic_cast<BitVectorVoxelList*>(other.get()))
    return ok(gvx_list_merge_list(m_h, l->handle(), off3(off, o), meaning), "merge");
  if (auto* m = dynamic_cast<voxelmap::BitVectorVoxelMap*>(other.get()))
    return ok(gvx_list_merge_map(m_h, m->handle(), off3(off, o), meaning), "merge");
  std::fprintf(stderr, "gpu_voxels: THIS OPERATION IS NOT YET SUPPORTED BY THE DATA STRUCTURE!\n");
  return false;
}

void BitVectorVoxelList::clearMap() { ok(gvx_list_clear(m_h), "clearMap"); }

std::size_t BitVectorVoxelList::collideWithBitcheck(const BitVectorVoxelList* other, const uint8_t margin,
                                                    const Vector3i& offset)
{
  if (offset.x || offset.y || offset.z) return 0;  // the reference refuses offsets here (BitVoxelList.hpp:398-402)
  uint64_t c = 0;
  if (!ok(gvx_list_collide_bitcheck(m_h, other->handle(), margin, &c), "collideWithBitcheck")) return SSIZE_MAX;
  return (std::size_t)c;
}

std::size_t BitVectorVoxelList::collideCountingPerMeaning(const GpuVoxelsMapSharedPtr other,
                                                          std::vector<std::size_t>& per_meaning, const Vector3i& offset)
{
  const BitVectorVoxelList* o = other ? other->as<BitVectorVoxelList>() : nullptr;
  if (!o || offset.x || offset.y || offset.z) return SSIZE_MAX;
  uint64_t counts[256], total = 0;
  if (!ok(gvx_list_collide_per_meaning(m_h, o->handle(), counts, &total), "collideCountingPerMeaning")) return SSIZE_MAX;
  if (per_meaning.size() < 256) per_meaning.resize(256, 0);
  for (int k = 0; k < 256; ++k) per_meaning[k] += (std::size_t)counts[k];
  return (std::size_t)total;
}

void BitVectorVoxelList::shiftLeftSweptVolumeIDs(uint8_t shift_size)
{
  ok(gvx_list_shift_sv(m_h, shift_size), "shiftLeftSweptVolumeIDs");
}

bool BitVectorVoxelList::writeToDisk(const std::string path) { return ok(gvx_list_write_file(m_h, path.c_str()), "writeToDisk"); }

bool BitVectorVoxelList::readFromDisk(const std::string path) { return ok(gvx_list_read_file(m_h, path.c_str()), "readFromDisk"); }

std::size_t BitVectorVoxelList::getMemoryUsage() const
{
  uint64_t n = 0;
  gvx_list_size(m_h, &n);
  return (std::size_t)(n * (4 + 12 + 32));
}

Vector3ui BitVectorVoxelList::getDimensions() const
{
  uint64_t n = 0;
  gvx_list_size(m_h, &n);
  return Vector3ui((uint32_t)n, 1, 1);
}

std::size_t BitVectorVoxelList::collideMap(gvx_map map, float thr, const Vector3i& off, int mode, const uint8_t* tm,
                                           BitVectorVoxel* types)
{
  int32_t o[3];
  uint64_t c = 0;
  uint8_t mask[32];
  int s = gvx_list_collide_map(m_h, map, thr, off3(off, o), mode, tm, &c, mask);
  if (s == GVX_ERR_DIM_MISMATCH) {
    warn("collideWith", s);
    return SSIZE_MAX;
  }
  if (!ok(s, "collideWith")) return SSIZE_MAX;
  if (types) {
    types->bitVector().clear();
    std::memcpy(types->bitVector().data(), mask, 32);
  }
  return (std::size_t)c;
}

std::size_t BitVectorVoxelList::collideWith(const voxelmap::BitVectorVoxelMap* m, float thr, const Vector3i& off)
{
  return collideMap(m->handle(), thr, off, 0, nullptr, nullptr);
}
std::size_t BitVectorVoxelList::collideWith(const voxelmap::ProbVoxelMap* m, float thr, const Vector3i& off)
{
  return collideMap(m->handle(), thr, off, 0, nullptr, nullptr);
}

// collideVoxellists (TemplateVoxelList.hpp:231-289): offsets are not supported there either
std::size_t BitVectorVoxelList::collideWith(const BitVectorVoxelList* other, float, const Vector3i& off)
{
  if (off != Vector3i(0)) {
    std::fprintf(stderr, "gpu_voxels: Offset for VoxelList collision not supported! Not performing collision check.\n");
    return 0;
  }
  uint64_t c = 0;
  if (!ok(gvx_list_collide_list(m_h, other->handle(), &c, nullptr), "collideWith")) return SSIZE_MAX;
  return (std::size_t)c;
}

std::size_t BitVectorVoxelList::collideWithTypes(const voxelmap::BitVectorVoxelMap* m, BitVectorVoxel& t, float thr,
                                                 const Vector3i& off)
{
  return collideMap(m->handle(), thr, off, 1, nullptr, &t);
}
std::size_t BitVectorVoxelList::collideWithTypes(const voxelmap::ProbVoxelMap* m, BitVectorVoxel& t, float thr,
                                                 const Vector3i& off)
{
  return collideMap(m->handle(), thr, off, 1, nullptr, &t);
}

std::size_t BitVectorVoxelList::collideWithTypes(const BitVectorVoxelList* other, BitVectorVoxel& t, float,
                                                 const Vector3i& off)
{
  if (off != Vector3i(0)) {
    std::fprintf(stderr, "gpu_voxels: Offset for VoxelList operation not supported! Result is undefined.\n");
    return 0;
  }
  uint64_t c = 0;
  uint8_t mask[32];
  if (!ok(gvx_list_collide_list(m_h, other->handle(), &c, mask), "collideWithTypes")) return SSIZE_MAX;
  std::memcpy(t.bitVector().data(), mask, 32);
  return (std::size_t)c;
}

std::size_t BitVectorVoxelList::collideWithTypeMask(const voxelmap::BitVectorVoxelMap* m, const BitVectorVoxel& types,
                                                    float thr, const Vector3i& off)
{
  return collideMap(m->handle(), thr, off, 2, types.bitVector().data(), nullptr);
}
std::size_t BitVectorVoxelList::collideWithTypeMask(const voxelmap::ProbVoxelMap* m, const BitVectorVoxel& types,
                                                    float thr, const Vector3i& off)
{
  return collideMap(m->handle(), thr, off, 2, types.bitVector().data(), nullptr);
}

void BitVectorVoxelList::copyCoordsToHost(std::vector<Vector3ui>& host_vec)
{
  uint64_t n = 0;
  gvx_list_size(m_h, &n);
  host_vec.resize(n);
  if (n) ok(gvx_list_read(m_h, nullptr, &host_vec[0].x, nullptr, GVX_MEM_HOST), "copyCoordsToHost");
}

void BitVectorVoxelList::copyVoxelsToHost(std::vector<uint32_t>& ids, std::vector<Vector3ui>& coords,
                                          std::vector<BitVectorVoxel>& voxels)
{
  uint64_t n = 0;
  gvx_list_size(m_h, &n);
  ids.resize(n);
  coords.resize(n);
  voxels.resize(n);
  if (n)
    ok(gvx_list_read(m_h, ids.data(), &coords[0].x, voxels[0].bitVector().data(), GVX_MEM_HOST), "copyVoxelsToHost");
}

uint32_t* BitVectorVoxelList::getDeviceIdPtr() const
{
  uint32_t* p = nullptr;
  gvx_list_device_ptrs(m_h, &p, nullptr, nullptr);
  return p;
}
Vector3ui* BitVectorVoxelList::getDeviceCoordPtr() const
{
  uint32_t* p = nullptr;
  gvx_list_device_ptrs(m_h, nullptr, &p, nullptr);
  return (Vector3ui*)p;
}
BitVectorVoxel* BitVectorVoxelList::getDeviceDataPtr() const
{
  uint8_t* p = nullptr;
  gvx_list_device_ptrs(m_h, nullptr, nullptr, &p);
  return (BitVectorVoxel*)p;
}

}  // namespace voxellist

// ---------------------------------------------------------------- robot
namespace robot {

void DHParameters::convertDHtoM(Matrix4f& m) const
{
  float p[5] = {theta, d, a, alpha, joint_type == PRISMATIC ? 1.0f : 0.0f};
  float r[12];
  gvx_dh_matrix(p, value, r);
  m = Matrix4f{r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8], r[9], r[10], r[11], 0, 0, 0, 1};
}

KinematicChain::KinematicChain(const std::vector<std::string>& linknames, const std::vector<DHParameters>& dh_params,
                               const MetaPointCloud& pointclouds, const Matrix4f& base)
    : m_names(linknames), m_dh(dh_params), m_clouds(pointclouds), m_base(base)
{
  if (m_dh.size() != m_names.size()) throw std::invalid_argument("one DH parameter set per link required");
}

KinematicChain::KinematicChain(const std::vector<std::string>& linknames, const std::vector<DHParameters>& dh_params,
                               const std::vector<std::string>& paths_to_pointclouds, const bool use_model_path)
    : KinematicChain(linknames, dh_params, MetaPointCloud(paths_to_pointclouds, paths_to_pointclouds, use_model_path))
{
}

void KinematicChain::setConfiguration(const JointValueMap& jointmap)
{
  for (std::size_t i = 0; i < m_names.size(); ++i) {
    auto it = jointmap.find(m_names[i]);
    if (it != jointmap.end()) m_dh[i].value = it->second;
  }
}

void KinematicChain::getConfiguration(JointValueMap& jointmap) const
{
  for (std::size_t i = 0; i < m_names.size(); ++i) jointmap[m_names[i]] = m_dh[i].value;
}

void KinematicChain::packed(std::vector<float>& dh, std::vector<float>& q, float base[12]) const
{
  dh.resize(5 * m_dh.size());
  q.resize(m_dh.size());
  for (std::size_t i = 0; i < m_dh.size(); ++i) {
    dh[5 * i] = m_dh[i].theta;
    dh[5 * i + 1] = m_dh[i].d;
    dh[5 * i + 2] = m_dh[i].a;
    dh[5 * i + 3] = m_dh[i].alpha;
    dh[5 * i + 4] = m_dh[i].joint_type == PRISMATIC ? 1.0f : 0.0f;
    q[i] = m_dh[i].value;
  }
  const Matrix4f& b = m_base;
  float t[12] = {b.a11, b.a12, b.a13, b.a14, b.a21, b.a22, b.a23, b.a24, b.a31, b.a32, b.a33, b.a34};
  std::memcpy(base, t, sizeof(t));
}

// Cloud i belongs to link i; links without a cloud contribute an empty cloud.
static std::vector<uint32_t> link_sizes(const MetaPointCloud& c, std::size_t n_links)
{
  std::vector<uint32_t> s = c.getPointcloudSizes();
  s.resize(n_links, 0);
  return s;
}

const MetaPointCloud* KinematicChain::getTransformedClouds()
{
  std::vector<float> dh, q;
  float base[12];
  packed(dh, q, base);
  std::vector<uint32_t> sizes = link_sizes(m_clouds, m_dh.size());
  const auto& p = m_clouds.points();
  std::vector<Vector3f> out(p.size());
  if (!p.empty())
    ok(gvx_robot_transform(&p[0].x, sizes.data(), (uint32_t)sizes.size(), dh.data(), base, q.data(), GVX_MEM_HOST,
                           &out[0].x, GVX_MEM_HOST),
       "getTransformedClouds");
  m_transformed = MetaPointCloud();
  std::size_t off = 0;
  for (uint32_t s : m_clouds.getPointcloudSizes()) {
    m_transformed.addCloud(std::vector<Vector3f>(out.begin() + off, out.begin() + off + s));
    off += s;
  }
  return &m_transformed;
}

bool KinematicChain::insertIntoMap(GpuVoxelsMap* map, BitVoxelMeaning meaning)
{
  std::vector<BitVoxelMeaning> ms(m_dh.size(), meaning);
  return insertIntoMap(map, ms);
}

bool KinematicChain::insertIntoMap(GpuVoxelsMap* map, const std::vector<BitVoxelMeaning>& meanings)
{
  std::vector<float> dh, q;
  float base[12];
  packed(dh, q, base);
  std::vector<uint32_t> sizes = link_sizes(m_clouds, m_dh.size()), mn = to_u32(meanings);
  mn.resize(sizes.size(), mn.empty() ? 0 : mn.back());
  const auto& p = m_clouds.points();
  const float* pp = p.empty() ? nullptr : &p[0].x;
  if (auto* vm = dynamic_cast<voxelmap::TemplateVoxelMapBase*>(map)) {
    int out = 0;
    return ok(gvx_map_insert_robot(vm->handle(), pp, sizes.data(), (uint32_t)sizes.size(), dh.data(), base, q.data(),
                                   mn.data(), 0, GVX_MEM_HOST, &out),
              "insertRobotIntoMap");
  }
  if (auto* vl = dynamic_cast<voxellist::BitVectorVoxelList*>(map))
    return ok(gvx_list_insert_robot(vl->handle(), pp, sizes.data(), (uint32_t)sizes.size(), dh.data(), base, q.data(),
                                    mn.data(), 0, GVX_MEM_HOST),
              "insertRobotIntoMap");
  return false;
}

bool KinematicChain::insertIntoMapSelfCollAware(GpuVoxelsMap* map, const std::vector<BitVoxelMeaning>& meanings,
                                                const std::vector<BitVector<BIT_VECTOR_LENGTH> >& masks,
                                                BitVector<BIT_VECTOR_LENGTH>* colliding)
{
  auto* bm = dynamic_cast<voxelmap::BitVectorVoxelMap*>(map);
  if (!bm) {
    const MetaPointCloud* posed = getTransformedClouds();
    return map->insertMetaPointCloudWithSelfCollisionCheck(posed, meanings, masks, colliding);
  }
  std::vector<float> dh, q;
  float base[12];
  packed(dh, q, base);
  std::vector<uint32_t> sizes = link_sizes(m_clouds, m_dh.size()), mn;
  std::vector<uint8_t> mk;
  if (!selfcoll_args(sizes.size(), meanings, masks, mn, mk)) return false;
  const auto& p = m_clouds.points();
  BitVector<BIT_VECTOR_LENGTH> col;
  int self = 0, out = 0;
  if (!ok(gvx_map_insert_robot_selfcoll(bm->handle(), p.empty() ? nullptr : &p[0].x, sizes.data(),
                                        (uint32_t)sizes.size(), dh.data(), base, q.data(),
                                        mn.empty() ? nullptr : mn.data(), mk.empty() ? nullptr : mk.data(),
                                        GVX_MEM_HOST, col.data(), &self, &out),
          "insertRobotIntoMapSelfCollAware"))
    return false;
  if (colliding) *colliding = col;
  return self != 0;
}

}  // namespace robot

// ---------------------------------------------------------------- facade
GpuVoxels::GpuVoxels() { ensure_init(); }
GpuVoxels::~GpuVoxels() {}

GpuVoxelsSharedPtr GpuVoxels::getInstance()
{
  static GpuVoxelsSharedPtr inst(new GpuVoxels());
  return inst;
}

void GpuVoxels::initialize(const uint32_t dx, const uint32_t dy, const uint32_t dz, const float side)
{
  if (m_initialized) {
    std::fprintf(stderr, "gpu_voxels: GpuVoxels already initialized, keeping the first dimensions\n");
    return;
  }
  m_dim = Vector3ui(dx, dy, dz);
  m_voxel_side_length = side;
  m_initialized = true;
}

GpuVoxelsMapSharedPtr GpuVoxels::addMap(const MapType map_type, const std::string& map_name)
{
  if (m_managed_maps.count(map_name)) {
    std::fprintf(stderr, "gpu_voxels: Map with name '%s' already exists.\n", map_name.c_str());
    return GpuVoxelsMapSharedPtr();
  }
  GpuVoxelsMapSharedPtr p;
  switch (map_type) {
    case MT_BITVECTOR_VOXELMAP: p.reset(new voxelmap::BitVectorVoxelMap(m_dim, m_voxel_side_length)); break;
    case MT_PROBAB_VOXELMAP: p.reset(new voxelmap::ProbVoxelMap(m_dim, m_voxel_side_length)); break;
    case MT_DISTANCE_VOXELMAP: p.reset(new voxelmap::DistanceVoxelMap(m_dim, m_voxel_side_length)); break;
    case MT_BITVECTOR_VOXELLIST: p.reset(new voxellist::BitVectorVoxelList(m_dim, m_voxel_side_length)); break;
    default: throw GPU_VOXELS_MAP_TYPE_NOT_IMPLEMENTED;
  }
  m_managed_maps[map_name] = p;
  return p;
}

bool GpuVoxels::delMap(const std::string& name) { return m_managed_maps.erase(name) > 0; }

bool GpuVoxels::clearMap(const std::string& name)
{
  auto it = m_managed_maps.find(name);
  if (it == m_managed_maps.end()) return false;
  it->second->clearMap();
  return true;
}

bool GpuVoxels::clearMap(const std::string& name, BitVoxelMeaning m)
{
  auto it = m_managed_maps.find(name);
  if (it == m_managed_maps.end()) return false;
  it->second->clearBitVoxelMeaning(m);
  return true;
}

GpuVoxelsMapSharedPtr GpuVoxels::getMap(const std::string& name)
{
  auto it = m_managed_maps.find(name);
  return it == m_managed_maps.end() ? GpuVoxelsMapSharedPtr() : it->second;
}

bool GpuVoxels::addRobot(const std::string& name, const std::vector<std::string>& links,
                         const std::vector<robot::DHParameters>& dh, const MetaPointCloud& clouds)
{
  if (m_managed_robots.count(name)) return false;
  m_managed_robots[name] = robot::RobotInterfaceSharedPtr(new robot::KinematicChain(links, dh, clouds));
  return true;
}

bool GpuVoxels::addRobot(const std::string& name, const std::vector<std::string>& links,
                         const std::vector<robot::DHParameters>& dh, const std::vector<std::string>& paths,
                         const bool use_model_path)
{
  if (m_managed_robots.count(name)) return false;
  m_managed_robots[name] = robot::RobotInterfaceSharedPtr(new robot::KinematicChain(links, dh, paths, use_model_path));
  return true;
}

bool GpuVoxels::insertPointCloudFromFile(const std::string map_name, const std::string path, const bool use_model_path,
                                         const BitVoxelMeaning voxel_meaning, const bool shift_to_zero,
                                         const Vector3f& offset_XYZ, const float scaling)
{
  GpuVoxelsMapSharedPtr m = getMap(map_name);
  if (!m) return false;
  return m->insertPointCloudFromFile(path, use_model_path, voxel_meaning, shift_to_zero, offset_XYZ, scaling);
}

robot::RobotInterfaceSharedPtr GpuVoxels::getRobot(const std::string& name)
{
  auto it = m_managed_robots.find(name);
  return it == m_managed_robots.end() ? robot::RobotInterfaceSharedPtr() : it->second;
}

bool GpuVoxels::setRobotConfiguration(std::string name, const robot::JointValueMap& jm)
{
  auto r = getRobot(name);
  if (!r) return false;
  r->setConfiguration(jm);
  return true;
}

bool GpuVoxels::getRobotConfiguration(const std::string& name, robot::JointValueMap& jm)
{
  auto r = getRobot(name);
  if (!r) return false;
  r->getConfiguration(jm);
  return true;
}

bool GpuVoxels::insertPointCloudIntoMap(const PointCloud& cloud, std::string name, const BitVoxelMeaning m)
{
  return insertPointCloudIntoMap(cloud.points(), name, m);
}

bool GpuVoxels::insertPointCloudIntoMap(const std::vector<Vector3f>& cloud, std::string name, const BitVoxelMeaning m)
{
  auto p = getMap(name);
  if (!p) return false;
  p->insertPointCloud(cloud, m);
  return true;
}

bool GpuVoxels::insertMetaPointCloudIntoMap(const MetaPointCloud& mpc, std::string name,
                                            const std::vector<BitVoxelMeaning>& ms)
{
  auto p = getMap(name);
  if (!p) return false;
  p->insertMetaPointCloud(mpc, ms);
  return true;
}

bool GpuVoxels::insertMetaPointCloudIntoMap(const MetaPointCloud& mpc, std::string name, const BitVoxelMeaning m)
{
  auto p = getMap(name);
  if (!p) return false;
  p->insertMetaPointCloud(mpc, m);
  return true;
}

// GpuVoxels.cpp:499-517: all transformed link clouds with one meaning
bool GpuVoxels::insertRobotIntoMap(std::string robot_name, std::string map_name, const BitVoxelMeaning m)
{
  auto r = getRobot(robot_name);
  auto p = getMap(map_name);
  if (!r || !p) return false;
  return r->insertIntoMap(p.get(), m);
}

bool GpuVoxels::insertRobotIntoMapSelfCollAware(std::string robot_name, std::string map_name,
                                                const std::vector<BitVoxelMeaning>& meanings,
                                                const std::vector<BitVector<BIT_VECTOR_LENGTH> >& masks,
                                                BitVector<BIT_VECTOR_LENGTH>* colliding)
{
  auto r = getRobot(robot_name);
  auto p = getMap(map_name);
  if (!r || !p) return false;
  return r->insertIntoMapSelfCollAware(p.get(), meanings, masks, colliding);
}

bool GpuVoxels::insertBoxIntoMap(const Vector3f& cmin, const Vector3f& cmax, std::string name, const BitVoxelMeaning m,
                                 uint16_t points_per_voxel)
{
  auto p = getMap(name);
  if (!p) return false;
  float delta = m_voxel_side_length / points_per_voxel;
  p->insertCoordinateList(geometry_generation::createBoxOfPoints(cmin, cmax, delta, m_voxel_side_length), m);
  return true;
}

void GpuVoxels::getDimensions(uint32_t& x, uint32_t& y, uint32_t& z)
{
  x = m_dim.x;
  y = m_dim.y;
  z = m_dim.z;
}
void GpuVoxels::getDimensions(Vector3ui& d) { d = m_dim; }
void GpuVoxels::getVoxelSideLength(float& s) { s = m_voxel_side_length; }

}  // namespace gpu_voxels
