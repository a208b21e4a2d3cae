// test_dropin.cpp -- the reference's hot-path tests, written against the C++ drop-in API.
//
// Each case restates a Boost.Test case of GPU-Voxels (packages/gpu_voxels/src/gpu_voxels/
// test/*.cu) with the same inputs and known answers, using gpu_voxels:: classes backed
// by libgvx.so.  Exit status = number of failed checks.
#include <gpu_voxels/GpuVoxels.h>

#include <algorithm>
#include <cstdio>
#include <sstream>
#include <string>

using namespace gpu_voxels;

static int g_fail = 0, g_pass = 0;
#define CHECK(cond, msg)                                                        \
  do {                                                                          \
    if (cond) { ++g_pass; }                                                     \
    else { ++g_fail; std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, msg); } \
  } while (0)

static const uint32_t dimX = 89, dimY = 123, dimZ = 74, numberOfPoints = 10000;  // testing_fixtures.hpp:137-179

// GeometryGeneration.cpp:237-299
static void createEquidistantPointsInBox(size_t max_n, Vector3ui c, float s, std::vector<Vector3f>& pts)
{
  for (uint32_t i = 0; i < (c.x - 1) / 2; i++)
    for (uint32_t j = 0; j < (c.y - 1) / 2; j++)
      for (uint32_t k = 0; k < (c.z - 1) / 2; k++) {
        if (pts.size() >= max_n) return;
        pts.push_back(Vector3f(i * 2 * s + s / 2.0, j * 2 * s + s / 2.0, k * 2 * s + s / 2.0));
      }
}
static void createCheckerboard(size_t max_n, Vector3ui c, float s, std::vector<Vector3f>& b, std::vector<Vector3f>& w)
{
  for (uint32_t i = 0; i < (c.x - 1) / 2; i++)
    for (uint32_t j = 0; j < (c.y - 1) / 2; j++)
      for (uint32_t k = 0; k < (c.z - 1) / 2; k++) {
        if (b.size() >= max_n) return;
        b.push_back(Vector3f(i * 2 * s + s / 2.0, j * 2 * s + s / 2.0, k * 2 * s + s / 2.0));
        w.push_back(Vector3f((i * 2 + 1) * s + s / 2.0, (j * 2 + 1) * s + s / 2.0, (k * 2 + 1) * s + s / 2.0));
      }
}

using geometry_generation::createBoxOfPoints;
using voxelmap::BitVectorVoxelMap;
using voxelmap::DistanceVoxelMap;
using voxelmap::ProbVoxelMap;
using voxellist::BitVectorVoxelList;

static void test_voxelmap()
{
  const Vector3ui dim(dimX, dimY, dimZ);
  {  // testing_voxelmap.cu:96-115 collision
    ProbVoxelMap m1(dim, 1.f), m2(dim, 1.f);
    std::vector<Vector3f> p;
    createEquidistantPointsInBox(numberOfPoints, dim, 1.f, p);
    m1.insertPointCloud(p, eBVM_OCCUPIED);
    m2.insertPointCloud(p, eBVM_OCCUPIED);
    CHECK(m1.collideWith(&m2, 0.1) == numberOfPoints, "All collisions detected.");
  }
  {  // testing_voxelmap.cu:119-142 collision_with_offset
    ProbVoxelMap m1(dim, 1.f), m2(dim, 1.f);
    m1.insertPointCloud(createBoxOfPoints(Vector3f(2.1, 2.1, 2.1), Vector3f(4.1, 4.1, 4.1), 0.5), eBVM_OCCUPIED);
    m2.insertPointCloud(createBoxOfPoints(Vector3f(3.1, 3.1, 3.1), Vector3f(5.1, 5.1, 5.1), 0.5), eBVM_OCCUPIED);
    CHECK(m1.collideWith(&m2, 0.1) == 8, "All collisions without offset detected.");
    CHECK(m1.collideWith(&m2, 0.1, Vector3i(-1, 0, -1)) == 18, "All collisions with negative offset detected.");
    CHECK(m2.collideWith(&m1, 0.1, Vector3i(1, 0, 1)) == 18, "All collisions with positive offset detected.");
  }
  {  // testing_voxelmap.cu:144-168 no_collision
    ProbVoxelMap m1(dim, 1.f), m2(dim, 1.f);
    std::vector<Vector3f> b, w;
    createCheckerboard(numberOfPoints, dim, 1.f, b, w);
    m1.insertPointCloud(b, eBVM_OCCUPIED);
    m2.insertPointCloud(w, eBVM_OCCUPIED);
    CHECK(m1.collideWith(&m2, 0.1) == 0, "No collisions detected.");
  }
  {  // testing_voxelmap.cu:313-360 collision_with_types_probab
    BitVectorVoxelMap m1(dim, 1.f);
    ProbVoxelMap m2(dim, 1.f);
    std::vector<Vector3f> p = {Vector3f(2, 2, 2), Vector3f(2, 2, 3), Vector3f(2, 3, 2), Vector3f(2, 3, 3),
                               Vector3f(3, 2, 2), Vector3f(3, 2, 3), Vector3f(3, 3, 2), Vector3f(3, 3, 3)};
    m1.insertPointCloud(p, BitVoxelMeaning(eBVM_SWEPT_VOLUME_START + 10));
    m1.insertPointCloud(p, BitVoxelMeaning(eBVM_SWEPT_VOLUME_START + 42));
    m1.insertPointCloud(p, BitVoxelMeaning(eBVM_SWEPT_VOLUME_START + 99));
    m2.insertPointCloud(p, eBVM_OCCUPIED);
    BitVectorVoxel types;
    size_t n = m1.collideWithTypes(&m2, types, ProbabilisticVoxel::probabilityToFloat(cSENSOR_MODEL_OCCUPIED));
    CHECK(n == p.size(), "8 collisions");
    CHECK(types.bitVector().getBit(14) && types.bitVector().getBit(46) && types.bitVector().getBit(103), "types");
    types.bitVector().clearBit(14);
    types.bitVector().clearBit(46);
    types.bitVector().clearBit(103);
    CHECK(types.bitVector().noneButEmpty(), "no other types");
  }
  {  // testing_voxelmap.cu:482-516 iostream_bitvoxel
    BitVectorVoxel v;
    for (BitVoxelMeaning m : {eBVM_FREE, eBVM_COLLISION, eBVM_UNKNOWN, eBVM_SWEPT_VOLUME_START, BitVoxelMeaning(111),
                              eBVM_SWEPT_VOLUME_END, eBVM_UNDEFINED})
      v.insert(m);
    std::string truth = std::string("10111") + std::string(106, '0') + "1" + std::string(142, '0') + "11";
    std::stringstream ss;
    ss << v;
    CHECK(ss.str() == truth, "Outstream");
  }
}

static void two_boxes(MetaPointCloud& mpc)
{
  std::vector<std::vector<Vector3f> > clouds;
  clouds.push_back(createBoxOfPoints(Vector3f(1.1, 1.1, 1.1), Vector3f(3.9, 3.9, 3.9), 0.1));
  clouds.push_back(createBoxOfPoints(Vector3f(2.1, 2.1, 2.1), Vector3f(4.9, 4.9, 4.9), 0.1));
  mpc = MetaPointCloud(clouds);
}

static void test_voxellist()
{
  const Vector3ui dim(dimX, dimY, dimZ);
  std::vector<BitVoxelMeaning> meanings = {BitVoxelMeaning(11), BitVoxelMeaning(12)};
  MetaPointCloud boxes;
  two_boxes(boxes);
  {  // testing_voxellist.cu:284-341
    BitVectorVoxelList list(dim, 1.f);
    list.insertMetaPointCloud(boxes, meanings);
    std::vector<uint32_t> ids;
    std::vector<Vector3ui> coords;
    std::vector<BitVectorVoxel> vox;
    list.copyVoxelsToHost(ids, coords, vox);
    CHECK(vox.size() == 46, "Number of reduced cubes == 46");
    BitVector<BIT_VECTOR_LENGTH> both;
    both.setBit(11);
    both.setBit(12);
    size_t nb = 0;
    for (size_t i = 0; i < vox.size(); ++i)
      if (vox[i].bitVector() == both) {
        ++nb;
        CHECK(coords[i].x >= 2 && coords[i].x <= 3 && coords[i].y >= 2 && coords[i].y <= 3 && coords[i].z >= 2 &&
                  coords[i].z <= 3,
              "merged voxel in the overlap");
      }
    CHECK(nb == 8, "Detect 8 overlapping voxels.");
  }
  {  // testing_voxellist.cu:111-148
    BitVectorVoxelList list(dim, 1.f);
    list.insertMetaPointCloud(boxes, meanings);
    ProbVoxelMap map(dim, 1.f);
    map.insertMetaPointCloud(boxes, eBVM_OCCUPIED);
    CHECK(list.collideWith(&map, ProbabilisticVoxel::probabilityToFloat(cSENSOR_MODEL_OCCUPIED)) == 46, "46");
  }
  for (int kind = 0; kind < 2; ++kind) {  // testing_voxellist.cu:151-238
    BitVectorVoxelList list(dim, 1.f);
    list.insertMetaPointCloud(boxes, meanings);
    BitVectorVoxel types;
    size_t n;
    if (kind == 0) {
      ProbVoxelMap map(dim, 1.f);
      map.insertMetaPointCloud(boxes, eBVM_OCCUPIED);
      n = list.collideWithTypes(&map, types, 1.0);
    } else {
      BitVectorVoxelMap map(dim, 1.f);
      map.insertMetaPointCloud(boxes, eBVM_OCCUPIED);
      n = list.collideWithTypes(&map, types, 1.0);
    }
    CHECK(n == 46, "Number of Collisions == 46");
    CHECK(types.bitVector().getBit(11) && types.bitVector().getBit(12), "Both Types found.");
    types.bitVector().clearBit(11);
    types.bitVector().clearBit(12);
    CHECK(types.bitVector().isZero(), "All other Types are clear.");
  }
  {  // testing_voxellist.cu:241-281
    BitVectorVoxelList list(dim, 1.f);
    std::vector<std::vector<Vector3f> > c(1, createBoxOfPoints(Vector3f(1.1, 1.1, 1.1), Vector3f(3.9, 3.9, 3.9), 0.1));
    MetaPointCloud box(c);
    list.insertMetaPointCloud(box, eBVM_OCCUPIED);
    ProbVoxelMap map(dim, 1.f);
    map.insertMetaPointCloud(box, eBVM_OCCUPIED);
    const float thr = ProbabilisticVoxel::probabilityToFloat(cSENSOR_MODEL_OCCUPIED);
    for (float shift = 0.0; shift < 4.0; shift += 0.5) {
      size_t n = list.collideWith(&map, thr, Vector3i(shift, 0, 0));
      size_t expect = shift < 1.0 ? 27 : shift < 2.0 ? 18 : shift < 3.0 ? 9 : 0;
      CHECK(n == expect, "shifted collisions");
    }
  }
  {  // testing_voxellist.cu:859-985 bitmasked_collision
    BitVectorVoxelList list(dim, 0.025);
    ProbVoxelMap map(dim, 0.025);
    map.insertPointCloud(createBoxOfPoints(Vector3f(0.111, 0.111, 0.111), Vector3f(0.799, 0.399, 0.399), 0.005),
                         eBVM_OCCUPIED);
    list.insertPointCloud(createBoxOfPoints(Vector3f(0.111, 0.311, 0.111), Vector3f(0.399, 0.599, 0.399), 0.005),
                          BitVoxelMeaning(34));
    list.insertPointCloud(createBoxOfPoints(Vector3f(0.311, 0.311, 0.111), Vector3f(0.599, 0.599, 0.399), 0.005),
                          BitVoxelMeaning(63));
    list.insertPointCloud(createBoxOfPoints(Vector3f(0.511, 0.311, 0.111), Vector3f(0.799, 0.599, 0.399), 0.005),
                          BitVoxelMeaning(102));
    const Vector3i offs[3] = {Vector3i(0, 0, 0), Vector3i(0, -4, 0), Vector3i(0, -4, -4)};
    const size_t expect[3][5] = {{0, 0, 9, 15, 21}, {0, 0, 18, 30, 42}, {0, 0, 12, 20, 28}};
    for (int o = 0; o < 3; ++o) {
      BitVectorVoxel t;
      CHECK(list.collideWithTypeMask(&map, t, 1.0f, offs[o]) == expect[o][0] * 64, "bitmask 0");
      const int bits[4] = {33, 34, 63, 102};
      for (int b = 0; b < 4; ++b) {
        t.bitVector().setBit(bits[b]);
        CHECK(list.collideWithTypeMask(&map, t, 1.0f, offs[o]) == expect[o][b + 1] * 64, "bitmask count");
      }
    }
  }
}

static void test_distance()
{
  // testing_distance.cu:39-127 distance_correctness: PBA == exactDistances3D
  const Vector3ui dim(128, 128, 131);
  std::vector<Vector3f> obst;
  createEquidistantPointsInBox(numberOfPoints, dim, 1.f, obst);
  DistanceVoxelMap ex(dim, 1.f), pba(dim, 1.f);
  ex.insertPointCloud(obst, eBVM_OCCUPIED);
  pba.insertPointCloud(obst, eBVM_OCCUPIED);
  pba.parallelBanding3D(1, 1, 1);
  ex.exactDistances3D(obst);
  std::vector<int32_t> a, b;
  pba.getSquaredDistances(a);
  ex.getSquaredDistances(b);
  CHECK(a == b, "PBA failed!");
  // testing_distance.cu:129-220 distance_extraction: 0 at obstacles, a free neighbour is farther
  Vector3ui o((uint32_t)obst[0].x, (uint32_t)obst[0].y, (uint32_t)obst[0].z);
  CHECK(pba.getSquaredObstacleDistance(o) == 0, "distance 0 at an obstacle");
  CHECK(pba.getSquaredObstacleDistance(o.x + 1, o.y, o.z) == 1, "neighbour at distance 1");
}

static void test_facade()
{
  GpuVoxelsSharedPtr gvl = GpuVoxels::getInstance();
  gvl->initialize(96, 96, 64, 0.02f);
  gvl->addMap(MT_PROBAB_VOXELMAP, "env");
  gvl->addMap(MT_BITVECTOR_VOXELMAP, "robot");
  gvl->addMap(MT_BITVECTOR_VOXELLIST, "robot_list");
  bool threw = false;
  try {
    gvl->addMap(MT_PROBAB_VOXELLIST, "nope");
  } catch (const std::string& e) {
    threw = (e == GPU_VOXELS_MAP_TYPE_NOT_IMPLEMENTED);
  }
  CHECK(threw, "unimplemented map type throws");
  CHECK(!gvl->addMap(MT_PROBAB_VOXELMAP, "env"), "duplicate name refused");
  CHECK(gvl->insertBoxIntoMap(Vector3f(0.6, 0.6, 0.0), Vector3f(1.2, 1.2, 0.4), "env", eBVM_OCCUPIED, 2), "box");
  std::vector<std::string> names = {"l0", "l1", "l2"};
  std::vector<robot::DHParameters> dh = {robot::DHParameters(0.2f, 0.f, 0.3f, 1.5707964f, 0.f, robot::REVOLUTE),
                                         robot::DHParameters(0.1f, 0.f, 0.3f, 0.f, 0.f, robot::REVOLUTE),
                                         robot::DHParameters(0.0f, 0.f, 0.2f, 0.f, 0.f, robot::REVOLUTE)};
  std::vector<std::vector<Vector3f> > clouds;
  for (int l = 0; l < 3; ++l) clouds.push_back(createBoxOfPoints(Vector3f(0, 0, 0), Vector3f(0.2, 0.04, 0.04), 0.01));
  CHECK(gvl->addRobot("arm", names, dh, MetaPointCloud(clouds, names)), "addRobot");
  auto rob = gvl->getRobot("arm");
  rob->setBaseTransformation(Matrix4f{1, 0, 0, 0.9, 0, 1, 0, 0.9, 0, 0, 1, 0.1, 0, 0, 0, 1});
  size_t hits = 0;
  for (float q = -3.1f; q < 3.1f; q += 0.3f) {
    robot::JointValueMap jm = {{"l0", q}, {"l1", 0.7f * q}, {"l2", -q}};
    gvl->setRobotConfiguration("arm", jm);
    gvl->clearMap("robot");
    gvl->clearMap("robot_list");
    CHECK(gvl->insertRobotIntoMap("arm", "robot", eBVM_OCCUPIED), "insertRobotIntoMap");
    CHECK(gvl->insertRobotIntoMap("arm", "robot_list", eBVM_OCCUPIED), "insertRobotIntoMap list");
    auto* env = gvl->getMap("env")->as<ProbVoxelMap>();
    size_t n_map = gvl->getMap("robot")->as<BitVectorVoxelMap>()->collideWith(env, 0.5);
    size_t n_list = gvl->getMap("robot_list")->as<BitVectorVoxelList>()->collideWith(env, 0.5);
    CHECK(n_map == n_list, "robot as map and as list collide identically");
    // transformed clouds inserted through the generic path give the same voxels
    ProbVoxelMap* dummy = nullptr;
    (void)dummy;
    BitVectorVoxelMap other(Vector3ui(96, 96, 64), 0.02f);
    other.insertMetaPointCloud(*rob->getTransformedClouds(), eBVM_OCCUPIED);
    CHECK(other.collideWith(env, 0.5) == n_map, "fused robot insert == transform + insertMetaPointCloud");
    hits += n_map;
  }
  CHECK(hits > 0, "the arm sweeps through the box");
}

// file formats: GPU_VOXELS_MODEL_PATH points at tests/golden/models (set by the runner)
static void test_files()
{
  using file_handling::PointcloudFileHandler;
  std::vector<Vector3f> pts;
  CHECK(PointcloudFileHandler::Instance()->loadPointCloud("arm_6_link.binvox", true, pts), "load binvox");
  CHECK(pts.size() == 250, "250 voxels in arm_6_link.binvox");
  std::vector<Vector3f> shifted;
  PointcloudFileHandler::Instance()->loadPointCloud("arm_6_link.binvox", true, shifted, true, Vector3f(0.1f, 0.2f, 0.3f));
  float mn[3] = {1e30f, 1e30f, 1e30f};
  for (const Vector3f& p : shifted) { mn[0] = std::min(mn[0], p.x); mn[1] = std::min(mn[1], p.y); mn[2] = std::min(mn[2], p.z); }
  CHECK(mn[0] == 0.1f && mn[1] == 0.2f && mn[2] == 0.3f, "shift to zero then offset");
  std::vector<Vector3f> none;
  CHECK(!PointcloudFileHandler::Instance()->loadPointCloud("no_such_file.binvox", true, none), "missing file fails");
  {
    voxelmap::BitVectorVoxelMap a(Vector3ui(64, 64, 64), 0.01f), b(Vector3ui(64, 64, 64), 0.01f);
    CHECK(a.insertPointCloudFromFile("arm_6_link.binvox", true, eBVM_OCCUPIED, true, Vector3f(0.1f, 0.1f, 0.1f)),
          "insertPointCloudFromFile");
    b.insertPointCloud(shifted, eBVM_OCCUPIED);  // same points, offset 0.1 0.2 0.3 -> not equal in y, z
    voxelmap::BitVectorVoxelMap c(Vector3ui(64, 64, 64), 0.01f);
    std::vector<Vector3f> same;
    PointcloudFileHandler::Instance()->loadPointCloud("arm_6_link.binvox", true, same, true, Vector3f(0.1f, 0.1f, 0.1f));
    c.insertPointCloud(same, eBVM_OCCUPIED);
    const std::size_t n_ac = a.collideWith(&c), n_cc = c.collideWith(&c);
    CHECK(n_ac > 0 && n_ac == n_cc, "file insert equals loaded insert");
    CHECK(a.writeToDisk("/tmp/gvx_dropin_map.dump"), "map writeToDisk");
    voxelmap::BitVectorVoxelMap d(Vector3ui(64, 64, 64), 0.01f);
    CHECK(d.readFromDisk("/tmp/gvx_dropin_map.dump"), "map readFromDisk");
    CHECK(d.collideWith(&c) == n_cc, "map dump round trip");
    voxelmap::ProbVoxelMap wrong(Vector3ui(64, 64, 64), 0.01f);
    CHECK(!wrong.readFromDisk("/tmp/gvx_dropin_map.dump"), "type mismatch refused");
    voxellist::BitVectorVoxelList l(Vector3ui(64, 64, 64), 0.01f), l2(Vector3ui(64, 64, 64), 0.01f);
    l.insertPointCloud(same, eBVM_OCCUPIED);
    CHECK(l.writeToDisk("/tmp/gvx_dropin_list.dump") && l2.readFromDisk("/tmp/gvx_dropin_list.dump"), "list dump");
    CHECK(l2.getDimensions().x == l.getDimensions().x && l2.collideWith(&c) == n_cc, "list dump round trip");
    std::remove("/tmp/gvx_dropin_map.dump");
    std::remove("/tmp/gvx_dropin_list.dump");
  }
  {
    // a two-link robot read from files (KinematicChain.cu:35-43)
    std::vector<std::string> links = {"arm_6_link.binvox", "arm_6_link.xyz"};
    std::vector<robot::DHParameters> dh = {robot::DHParameters(0.0f, 0.0f, 0.1f, 0.0f, 0.0f, robot::REVOLUTE),
                                           robot::DHParameters(0.0f, 0.0f, 0.1f, 0.0f, 0.0f, robot::REVOLUTE)};
    robot::KinematicChain chain(links, dh, links, true);
    CHECK(chain.getNumberOfLinks() == 2, "robot from files");
    voxelmap::BitVectorVoxelMap m(Vector3ui(128, 128, 128), 0.01f);
    chain.setBaseTransformation(Matrix4f{1, 0, 0, 0.6f, 0, 1, 0, 0.6f, 0, 0, 1, 0.6f, 0, 0, 0, 1});
    CHECK(chain.insertIntoMap(&m, eBVM_OCCUPIED), "insert file robot");
    CHECK(m.collideWith(&m) > 0, "file robot occupies voxels");
  }
}

// insertSensorData: a ray along +x marks the cells before the hit free, the hit occupied
static void test_sensor()
{
  voxelmap::ProbVoxelMap m(Vector3ui(16, 8, 8), 1.0f);
  std::vector<Vector3f> p = {Vector3f(6.2f, 2.9f, 3.1f)};
  m.insertSensorData(PointCloud(p), Vector3f(1.5f, 2.5f, 3.5f), true, false, eBVM_OCCUPIED, nullptr);
  std::vector<int8_t> v(16 * 8 * 8);
  gvx_map_read(m.handle(), v.data(), GVX_MEM_HOST);
  const uint32_t row = (3 * 8 + 2) * 16;
  bool ok_ray = v[row + 6] == 62;
  for (uint32_t x = 1; x < 6; ++x) ok_ray = ok_ray && v[row + x] == -10;
  CHECK(ok_ray, "ray: free cells then the occupied end voxel");
  voxelmap::BitVectorVoxelMap robot(Vector3ui(16, 8, 8), 1.0f);
  robot.insertPointCloud(p, (BitVoxelMeaning)11);
  voxelmap::ProbVoxelMap m2(Vector3ui(16, 8, 8), 1.0f);
  m2.insertSensorData(PointCloud(p), Vector3f(1.5f, 2.5f, 3.5f), true, true, (BitVoxelMeaning)11,
                      robot.getDeviceDataPtr());
  gvx_map_read(m2.handle(), v.data(), GVX_MEM_HOST);
  bool untouched = true;
  for (int8_t b : v) untouched = untouched && b == -128;
  CHECK(untouched, "points on the robot are cut");
}

// swept volumes: testing_bitvector.cu:144-196, testing_voxelmap.cu:244-299,
// testing_voxellist.cu:46-101 and 502-546
static void test_swept_volume()
{
  for (uint8_t shift = 3; shift < 20; shift += 10) {
    BitVector<BIT_VECTOR_LENGTH> v;
    for (uint32_t b : {0u, 1u, 2u, 3u, 4u, 14u, 24u, 254u, 255u}) v.setBit(b);
    performLeftShift(v, shift);
    bool gone = true;
    for (uint32_t b : {0u, 1u, 2u, 3u, 4u, 14u, 24u, 254u, 255u}) gone = gone && !v.getBit(b);
    CHECK(gone, "original bits cleared");
    CHECK(v.getBit(24 - shift) == (24 - shift > 3) && v.getBit(254 - shift) && v.getBit(255 - shift), "shifted bits");
  }
  std::vector<Vector3f> pts = {Vector3f(0.3f, 0.3f, 0.3f), Vector3f(1.3f, 1.3f, 1.3f), Vector3f(2.3f, 2.3f, 2.3f),
                               Vector3f(3.3f, 3.3f, 3.3f)};
  voxelmap::BitVectorVoxelMap m1(Vector3ui(dimX, dimY, dimZ), 1.f), m2(Vector3ui(dimX, dimY, dimZ), 1.f);
  bool all4 = true;
  for (uint32_t shift = 0; shift < 49; ++shift) {
    m1.clearMap();
    m2.clearMap();
    m2.insertPointCloud(pts, BitVoxelMeaning(54));
    m1.insertPointCloud(pts, BitVoxelMeaning(54 - shift));
    m2.shiftLeftSweptVolumeIDs(shift);
    BitVectorVoxel types;
    all4 = all4 && m1.collisionCheckBitvector(&m2, SVCollider(1.0, 1), types.bitVector()) == 4;
  }
  CHECK(all4, "map bit-shift collisions (window 1)");
  std::vector<Vector3f> one = {Vector3f(0.3f, 0.3f, 0.3f)};
  bool all1 = true;
  for (uint32_t shift = 0; shift < 49; ++shift) {
    voxellist::BitVectorVoxelList l1(Vector3ui(dimX, dimY, dimZ), 1.f), l2(Vector3ui(dimX, dimY, dimZ), 1.f);
    l2.insertPointCloud(one, BitVoxelMeaning(54));
    l1.insertPointCloud(one, BitVoxelMeaning(54 - shift));
    l2.shiftLeftSweptVolumeIDs(shift);
    all1 = all1 && l1.collideWithBitcheck(&l2, 1) == 1;
  }
  CHECK(all1, "list bit-shift collisions (window 1)");
  auto* obstacle = new voxellist::BitVectorVoxelList(Vector3ui(dimX, dimY, dimZ), 1.f);
  GpuVoxelsMapSharedPtr obstacle_ptr(obstacle);
  std::vector<std::vector<Vector3f> > box = {geometry_generation::createBoxOfPoints(Vector3f(8, 8, 8), Vector3f(11, 11, 11), 0.1f)};
  obstacle->insertMetaPointCloud(MetaPointCloud(box), eBVM_OCCUPIED);
  voxellist::BitVectorVoxelList swept(Vector3ui(dimX, dimY, dimZ), 1.f);
  std::vector<std::vector<Vector3f> > clouds;
  std::vector<BitVoxelMeaning> meanings;
  const float ratio = 0.02f;
  for (int i = 0; i < 50; i++) {
    Vector3f p((30 - 0) * ratio * i + 0, (30 - 0) * ratio * i + 0, (30 - 0) * ratio * i + 0);
    clouds.push_back(geometry_generation::createBoxOfPoints(p, Vector3f(p.x + 2.5f, p.y + 2.5f, p.z + 2.5f), 0.1f));
    meanings.push_back(BitVoxelMeaning(eBVM_SWEPT_VOLUME_START + i));
  }
  swept.insertMetaPointCloud(MetaPointCloud(clouds), meanings);
  std::vector<std::size_t> per(BIT_VECTOR_LENGTH, 0);
  const std::size_t total = swept.collideCountingPerMeaning(obstacle_ptr, per);
  std::stringstream ss;
  ss << "(Meaning|Collisions) ";
  for (std::size_t i = 0; i < BIT_VECTOR_LENGTH; i++)
    if (per[i] > 0) ss << "(" << i << "|" << per[i] << ") ";
  CHECK(ss.str() == "(Meaning|Collisions) (14|1) (15|8) (16|8) (17|27) (18|27) (19|27) (20|8) (21|1) (22|1) ",
        "per-meaning swept-volume collisions");
  CHECK(total == 108, "collisions == 108");
}

// test/testing_voxelmap.cu:366-479 through the drop-in classes, plus the robot form
static void test_selfcollision()
{
  using geometry_generation::createBoxOfPoints;
  for (int j = 2; j < 11; j += 8) {
    voxelmap::BitVectorVoxelMap map(Vector3ui(dimX, dimY, dimZ), 1.f);
    std::vector<std::vector<Vector3f> > col = {
        createBoxOfPoints(Vector3f(2.1, 2.1, 2.1), Vector3f(5.1, 5.1, 5.1), 1.0 / j),
        createBoxOfPoints(Vector3f(4.1, 4.1, 4.1), Vector3f(7.1, 7.1, 7.1), 1.0 / j),
        createBoxOfPoints(Vector3f(6.1, 6.1, 6.1), Vector3f(9.1, 9.1, 9.1), 1.0 / j),
        createBoxOfPoints(Vector3f(10.1, 10.1, 10.1), Vector3f(12.1, 12.1, 12.1), 1.0 / j)};
    std::vector<std::vector<Vector3f> > sep = {
        createBoxOfPoints(Vector3f(30.1, 30.1, 30.1), Vector3f(32.1, 32.1, 32.1), 1.0 / j),
        createBoxOfPoints(Vector3f(40.1, 40.1, 40.1), Vector3f(43.1, 43.1, 43.1), 1.0 / j),
        createBoxOfPoints(Vector3f(60.1, 60.1, 60.1), Vector3f(64.1, 64.1, 64.1), 1.0 / j),
        createBoxOfPoints(Vector3f(65.1, 65.1, 65.1), Vector3f(68.1, 68.1, 68.1), 1.0 / j)};
    MetaPointCloud mc(col), ms(sep);
    CHECK(map.insertMetaPointCloudWithSelfCollisionCheck(&mc) == true, "self collision detected");
    CHECK(map.insertMetaPointCloudWithSelfCollisionCheck(&ms) == false, "no self collision");
    map.clearMap();
    std::vector<BitVoxelMeaning> meanings = {BitVoxelMeaning(51), BitVoxelMeaning(52), BitVoxelMeaning(53),
                                             BitVoxelMeaning(54)};
    BitVector<BIT_VECTOR_LENGTH> all;
    all = ~all;
    all.clearBit(eBVM_COLLISION);
    std::vector<BitVector<BIT_VECTOR_LENGTH> > masks(4, all);
    masks[0].clearBit(52);
    masks[1].clearBit(51);
    for (int k = 0; k < 4; ++k) masks[k].clearBit(51 + k);
    BitVector<BIT_VECTOR_LENGTH> colliding;
    CHECK(map.insertMetaPointCloudWithSelfCollisionCheck(&mc, meanings, masks, &colliding) == true, "masked: collision");
    CHECK(!colliding.getBit(51) && colliding.getBit(52) && colliding.getBit(53) && !colliding.getBit(54),
          "colliding meanings 52 and 53");
    colliding.clearBit(52);
    colliding.clearBit(53);
    CHECK(colliding.isZero(), "nothing else collides");
    for (int k = 0; k < 4; ++k) {
      masks[k] = all;
      masks[k].clearBit(51 + k);
    }
    colliding.clear();
    CHECK(map.insertMetaPointCloudWithSelfCollisionCheck(&ms, meanings, masks, &colliding) == false && colliding.isZero(),
          "separated clouds do not collide");
    voxelmap::ProbVoxelMap prob(Vector3ui(dimX, dimY, dimZ), 1.f);
    CHECK(prob.insertMetaPointCloudWithSelfCollisionCheck(&mc) == true, "unsupported map type returns true");
  }
  // a two-link robot whose second link folds back onto the first
  GpuVoxelsSharedPtr gvl = GpuVoxels::getInstance();  // initialized by test_facade
  gvl->addMap(MT_BITVECTOR_VOXELMAP, "selfcoll");
  std::vector<robot::DHParameters> dh(2);
  std::vector<std::vector<Vector3f> > links = {
      createBoxOfPoints(Vector3f(0.3, 0.3, 0.3), Vector3f(0.5, 0.5, 0.5), 0.01),
      createBoxOfPoints(Vector3f(0.4, 0.4, 0.4), Vector3f(0.6, 0.6, 0.6), 0.01)};
  CHECK(gvl->addRobot("folding", {"l0", "l1"}, dh, MetaPointCloud(links)), "robot added");
  robot::JointValueMap q;
  q["l0"] = 0.f;
  q["l1"] = 0.f;
  gvl->setRobotConfiguration("folding", q);
  BitVector<BIT_VECTOR_LENGTH> robcol;
  CHECK(gvl->insertRobotIntoMapSelfCollAware("folding", "selfcoll", std::vector<BitVoxelMeaning>(),
                                             std::vector<BitVector<BIT_VECTOR_LENGTH> >(), &robcol),
        "overlapping links collide");
  CHECK(robcol.getBit(eBVM_SWEPT_VOLUME_START) && robcol.getBit(eBVM_SWEPT_VOLUME_START + 1), "both link meanings");
  gvl->delMap("selfcoll");
}

int main()
{
  test_voxelmap();
  test_voxellist();
  test_distance();
  test_facade();
  test_files();
  test_sensor();
  test_swept_volume();
  test_selfcollision();
  std::printf("drop-in tests: %d passed, %d failed\n", g_pass, g_fail);
  return g_fail;
}
