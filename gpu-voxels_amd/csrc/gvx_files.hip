// gvx_files.hip -- host-side file formats on the path: map and voxel-list dumps
// (TemplateVoxelMap / TemplateVoxelList writeToDisk, readFromDisk) and the point
// cloud readers behind insertPointCloudFromFile and file-based robots
// (PointcloudFileHandler with its xyz, PCD and binvox readers).  No GPU kernels:
// dumps stream the device data through a bounded host buffer.
#include "gvx_internal.hpp"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace gvx {

int list_assign_host(gvx_list l, uint64_t n, const uint32_t* ids, const uint32_t* coords, const uint8_t* bits);

constexpr size_t kChunk = 64u << 20;  // host staging per copy when dumping maps

// a * b without wrap-around: false when the product does not fit 64 bits
static bool mul_ok(uint64_t a, uint64_t b, uint64_t& out)
{
  if (a != 0 && b > UINT64_MAX / a) return false;
  out = a * b;
  return true;
}

// bytes between the stream's position and its end (0 when it cannot be told)
static uint64_t bytes_left(std::istream& in)
{
  const std::streampos here = in.tellg();
  if (here < 0) return 0;
  in.seekg(0, std::ios::end);
  const std::streampos end = in.tellg();
  in.seekg(here);
  return end > here ? (uint64_t)(end - here) : 0;
}

// ---------------------------------------------------------------- point clouds
// XyzFileReader.cpp:29-56: every line is read as whitespace separated numbers,
// consumed three at a time; a line's trailing one or two numbers are dropped.
static bool read_xyz(const std::string& path, std::vector<float>& pts)
{
  std::ifstream f(path.c_str());
  if (!f) return false;
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream iss(line);
    float x, y, z;
    while (iss >> x && iss >> y && iss >> z) {
      pts.push_back(x);
      pts.push_back(y);
      pts.push_back(z);
    }
  }
  return true;
}

// BinvoxFileReader.cpp:42-163: "#binvox <version>", then keywords up to "data"
// (dim d h w / translate tx ty tz / scale s, unknown keywords skip their line),
// one linefeed byte, then (value, count) byte pairs run-length coding d*h*w
// voxels with y fastest, then z, then x.  Voxels equal to 1 become the points
// (s/w * x + tx, s/w * y + ty, s/w * z + tz).
static bool read_binvox(const std::string& path, std::vector<float>& pts)
{
  std::ifstream in(path.c_str(), std::ios::in | std::ios::binary);
  if (!in.good()) return false;
  std::string tok;
  in >> tok;
  if (tok != "#binvox") return false;
  int version = 0;
  in >> version;
  int depth = -1, height = 0, width = 0;
  float tx = 0.f, ty = 0.f, tz = 0.f, scale = 0.f;
  bool done = false;
  while (in.good() && !done) {
    in >> tok;
    if (tok == "data") {
      done = true;
    } else if (tok == "dim") {
      in >> depth >> height >> width;
    } else if (tok == "translate") {
      in >> tx >> ty >> tz;
    } else if (tok == "scale") {
      in >> scale;
      scale = scale / (float)width;
    } else {
      int c;
      do {
        c = in.get();
      } while (in.good() && c != '\n');
    }
  }
  if (!done || depth < 0 || height < 0 || width < 0) return false;
  const uint64_t size = (uint64_t)depth * (uint64_t)height * (uint64_t)width;  // < 2^93: checked below
  in.get();  // the linefeed after "data"
  // every (value, count) pair covers at most 255 voxels: a grid the data cannot cover is malformed
  uint64_t cover;
  if (!mul_ok(bytes_left(in) / 2, 255, cover) || size > cover) return false;
  std::vector<uint8_t> vox(size, 0);
  uint64_t index = 0;
  while (index < size) {
    const int v = in.get();
    const int c = in.get();
    if (!in.good()) break;
    const uint64_t end = index + (uint64_t)(uint8_t)c;
    if (end > size) return false;
    std::fill(vox.begin() + index, vox.begin() + end, (uint8_t)v);
    index = end;
  }
  uint64_t i = 0;
  for (int x = 0; x < depth; ++x)
    for (int z = 0; z < height; ++z)
      for (int y = 0; y < width; ++y, ++i)
        if (vox[i] == 1) {
          pts.push_back(scale * (float)x + tx);
          pts.push_back(scale * (float)y + ty);
          pts.push_back(scale * (float)z + tz);
        }
  return true;
}

// LZF decompression (the block format of liblzf, which PCD's binary_compressed
// uses): a control byte c < 32 starts a literal run of c + 1 bytes; otherwise
// the top three bits are a length (7 = read one more byte and add it), the low
// five bits and the next byte an offset: copy length + 2 bytes from
// out - offset - 1.
static bool lzf_decompress(const uint8_t* in, size_t n_in, uint8_t* out, size_t n_out)
{
  size_t ip = 0, op = 0;
  while (ip < n_in) {
    unsigned c = in[ip++];
    if (c < 32) {
      const size_t len = c + 1;
      if (ip + len > n_in || op + len > n_out) return false;
      std::memcpy(out + op, in + ip, len);
      ip += len;
      op += len;
    } else {
      size_t len = c >> 5;
      if (len == 7) {
        if (ip >= n_in) return false;
        len += in[ip++];
      }
      if (ip >= n_in) return false;
      const size_t back = ((size_t)(c & 0x1f) << 8) + in[ip++] + 1;
      len += 2;
      if (back > op || op + len > n_out) return false;
      for (size_t k = 0; k < len; ++k, ++op) out[op] = out[op - back];  // may overlap
    }
  }
  return op == n_out;
}

static bool pcd_value(const uint8_t* p, char type, int size, float& out)
{
  switch (type) {
    case 'F':
      if (size == 4) { float v; std::memcpy(&v, p, 4); out = v; return true; }
      if (size == 8) { double v; std::memcpy(&v, p, 8); out = (float)v; return true; }
      return false;
    case 'I':
      if (size == 1) { out = (float)*(const int8_t*)p; return true; }
      if (size == 2) { int16_t v; std::memcpy(&v, p, 2); out = (float)v; return true; }
      if (size == 4) { int32_t v; std::memcpy(&v, p, 4); out = (float)v; return true; }
      if (size == 8) { int64_t v; std::memcpy(&v, p, 8); out = (float)v; return true; }
      return false;
    case 'U':
      if (size == 1) { out = (float)*p; return true; }
      if (size == 2) { uint16_t v; std::memcpy(&v, p, 2); out = (float)v; return true; }
      if (size == 4) { uint32_t v; std::memcpy(&v, p, 4); out = (float)v; return true; }
      if (size == 8) { uint64_t v; std::memcpy(&v, p, 8); out = (float)v; return true; }
      return false;
  }
  return false;
}

// PcdFileReader.cpp:31-54 loads a PCD file as pcl::PointXYZ (x, y, z of every
// point, NaN points kept).  The PCD format: header lines VERSION, FIELDS, SIZE,
// TYPE, COUNT, WIDTH, HEIGHT, VIEWPOINT, POINTS, DATA {ascii | binary |
// binary_compressed}; ascii = one point per line, binary = packed points,
// binary_compressed = two uint32 sizes + LZF data holding the fields one after
// the other (all points of field 0, then field 1, ...).
static bool read_pcd(const std::string& path, std::vector<float>& pts)
{
  std::ifstream in(path.c_str(), std::ios::in | std::ios::binary);
  if (!in.good()) return false;
  std::vector<std::string> fields;
  std::vector<int> sizes, counts;
  std::vector<char> types;
  uint64_t width = 0, height = 1, npoints = 0;
  bool have_points = false;
  std::string data, line;
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line.empty() || line[0] == '#') continue;
    std::istringstream ls(line);
    std::string key;
    ls >> key;
    if (key == "FIELDS" || key == "COLUMNS") {
      std::string f;
      while (ls >> f) fields.push_back(f);
    } else if (key == "SIZE") {
      int v;
      while (ls >> v) sizes.push_back(v);
    } else if (key == "TYPE") {
      std::string t;
      while (ls >> t) types.push_back(t.empty() ? '?' : t[0]);
    } else if (key == "COUNT") {
      int v;
      while (ls >> v) counts.push_back(v);
    } else if (key == "WIDTH") {
      ls >> width;
    } else if (key == "HEIGHT") {
      ls >> height;
    } else if (key == "POINTS") {
      ls >> npoints;
      have_points = true;
    } else if (key == "DATA") {
      ls >> data;
      break;
    }
  }
  const size_t nf = fields.size();
  if (nf == 0 || sizes.size() != nf || types.size() != nf) return false;
  if (counts.empty()) counts.assign(nf, 1);
  if (counts.size() != nf) return false;
  uint64_t n = npoints;
  if (!have_points && !mul_ok(width, height, n)) return false;
  int fx = -1, fy = -1, fz = -1;
  std::vector<size_t> off(nf, 0);
  uint64_t stride = 0;
  for (size_t i = 0; i < nf; ++i) {
    // a field is SIZE bytes (1, 2, 4 or 8) times COUNT (>= 1) values
    if (!(sizes[i] == 1 || sizes[i] == 2 || sizes[i] == 4 || sizes[i] == 8) || counts[i] < 1) return false;
    off[i] = (size_t)stride;
    uint64_t fb;
    if (!mul_ok((uint64_t)sizes[i], (uint64_t)counts[i], fb) || stride + fb < stride) return false;
    stride += fb;
    if (fields[i] == "x") fx = (int)i;
    if (fields[i] == "y") fy = (int)i;
    if (fields[i] == "z") fz = (int)i;
  }
  if (fx < 0 || fy < 0 || fz < 0) return false;
  const int fxyz[3] = {fx, fy, fz};
  // the point data follows the header: n points cannot need more bytes than the file has left
  // (ascii: at least two characters per point; binary: n * stride; compressed: the sizes below)
  const uint64_t left = bytes_left(in);
  uint64_t need;
  if (!mul_ok(n, stride, need)) return false;
  if (data == "ascii" && n > left / 2) return false;
  if (data == "binary" && need > left) return false;
  pts.reserve(pts.size() + 3 * (size_t)std::min<uint64_t>(n, left));
  if (data == "ascii") {
    // one point per line, every field's COUNT values in order
    size_t idx[3];
    for (int a = 0; a < 3; ++a) {
      size_t k = 0;
      for (int i = 0; i < fxyz[a]; ++i) k += (size_t)counts[i];
      idx[a] = k;
    }
    uint64_t got = 0;
    while (got < n && std::getline(in, line)) {
      std::istringstream ls(line);
      std::vector<std::string> tok;
      std::string t;
      while (ls >> t) tok.push_back(t);
      if (tok.empty()) continue;
      for (int a = 0; a < 3; ++a) {
        if (idx[a] >= tok.size()) return false;
        const std::string& s = tok[idx[a]];
        float v;
        if (s == "nan" || s == "NaN" || s == "NAN") v = std::nanf("");
        else if (types[fxyz[a]] == 'F') v = std::strtof(s.c_str(), nullptr);
        else v = (float)std::strtod(s.c_str(), nullptr);
        pts.push_back(v);
      }
      ++got;
    }
    return got == n;
  }
  std::vector<uint8_t> raw;
  if (data == "binary") {
    raw.resize((size_t)need);
    in.read((char*)raw.data(), (std::streamsize)raw.size());
    if ((uint64_t)in.gcount() != raw.size()) return false;
    for (uint64_t p = 0; p < n; ++p)
      for (int a = 0; a < 3; ++a) {
        float v;
        if (!pcd_value(raw.data() + p * stride + off[fxyz[a]], types[fxyz[a]], sizes[fxyz[a]], v)) return false;
        pts.push_back(v);
      }
    return true;
  }
  if (data == "binary_compressed") {
    uint32_t csize = 0, usize = 0;
    in.read((char*)&csize, 4);
    in.read((char*)&usize, 4);
    // an LZF back-reference of 3 input bytes yields at most 264 output bytes: the compressed size
    // bounds the data
    if (!in.good() || usize != need || csize > left || need > 88 * (uint64_t)csize + 264) return false;
    std::vector<uint8_t> comp(csize);
    in.read((char*)comp.data(), csize);
    if ((uint32_t)in.gcount() != csize) return false;
    raw.resize(usize);
    if (usize && !lzf_decompress(comp.data(), csize, raw.data(), usize)) return false;
    for (uint64_t p = 0; p < n; ++p)
      for (int a = 0; a < 3; ++a) {
        const int f = fxyz[a];
        float v;
        if (!pcd_value(raw.data() + off[f] * n + p * (size_t)sizes[f] * counts[f], types[f], sizes[f], v)) return false;
        pts.push_back(v);
      }
    return true;
  }
  return false;
}

// getGpuVoxelsPath (common_defines.h:285-300) joined with the file name the
// way boost::filesystem's operator/ joins them
static std::string model_path(const char* path, int use_model_path)
{
  std::string p(path);
  if (!use_model_path) return p;
  const char* env = std::getenv("GPU_VOXELS_MODEL_PATH");
  if (!env || !*env) return p;
  std::string b(env);
  if (p.empty()) return b;
  if (b.back() == '/' || p.front() == '/') return b + p;
  return b + "/" + p;
}

}  // namespace gvx

using namespace gvx;

extern "C" {

int gvx_load_point_cloud(const char* path, int use_model_path, int shift_to_zero, const float offset[3], float scaling,
                         float** xyz, uint64_t* n)
{
  if (!path || !xyz || !n) return fail(GVX_ERR_BAD_ARG, "gvx_load_point_cloud: null argument");
  *xyz = nullptr;
  *n = 0;
  const std::string p = model_path(path, use_model_path);
  std::vector<float> pts;
  bool ok;
  try {
    // PointcloudFileHandler.cpp:82-125: the format is chosen by the first of
    // "xyz", "pcd", "binvox" found anywhere in the path
    if (p.find("xyz") != std::string::npos) ok = read_xyz(p, pts);
    else if (p.find("pcd") != std::string::npos) ok = read_pcd(p, pts);
    else if (p.find("binvox") != std::string::npos) ok = read_binvox(p, pts);
    else return fail(GVX_ERR_UNSUPPORTED, p + " has no known file format");
  } catch (const std::bad_alloc&) {
    return fail(GVX_ERR_OOM, "out of host memory reading " + p);
  } catch (const std::exception& e) {
    return fail(GVX_ERR_BAD_ARG, "could not read point cloud " + p + ": " + e.what());
  }
  if (!ok) return fail(GVX_ERR_BAD_ARG, "could not read point cloud " + p);
  const uint64_t np = pts.size() / 3;
  if (shift_to_zero && np) {  // shiftPointCloudToZero (PointcloudFileHandler.cpp:177-194)
    float mn[3] = {pts[0], pts[1], pts[2]};
    for (uint64_t i = 1; i < np; ++i)
      for (int a = 0; a < 3; ++a) mn[a] = std::min(mn[a], pts[3 * i + a]);
    for (uint64_t i = 0; i < np; ++i)
      for (int a = 0; a < 3; ++a) pts[3 * i + a] -= mn[a];
  }
  const float o[3] = {offset ? offset[0] : 0.f, offset ? offset[1] : 0.f, offset ? offset[2] : 0.f};
  for (uint64_t i = 0; i < np; ++i)
    for (int a = 0; a < 3; ++a) pts[3 * i + a] = (scaling * pts[3 * i + a]) + o[a];
  float* out = (float*)std::malloc(np ? np * 3 * sizeof(float) : sizeof(float));
  if (!out) return fail(GVX_ERR_OOM, "gvx_load_point_cloud: host allocation failed");
  if (np) std::memcpy(out, pts.data(), np * 3 * sizeof(float));
  *xyz = out;
  *n = np;
  return GVX_OK;
}

void gvx_free(void* p) { std::free(p); }

static int file_map_write_file(gvx_map m, const char* path)
{
  if (!m || !path) return fail(GVX_ERR_BAD_ARG, "gvx_map_write_file: null argument");
  std::lock_guard<std::recursive_mutex> g(m->mu);
  std::ofstream out(path, std::ios::out | std::ios::binary | std::ios::trunc);
  if (!out.is_open()) return fail(GVX_ERR_BAD_ARG, std::string("cannot write ") + path);
  const int32_t type = m->type;
  out.write((const char*)&type, 4);
  out.write((const char*)&m->side, 4);
  out.write((const char*)m->dims, 12);
  std::vector<char> buf(std::min<uint64_t>(m->bytes, kChunk));
  for (uint64_t o = 0; o < m->bytes; o += buf.size()) {
    const size_t len = (size_t)std::min<uint64_t>(buf.size(), m->bytes - o);
    GVX_CHECK(copy_out(buf.data(), (const char*)m->data + o, len, GVX_MEM_HOST));
    GVX_CHECK(sync());
    out.write(buf.data(), (std::streamsize)len);
  }
  out.close();
  if (!out) return fail(GVX_ERR_BAD_ARG, std::string("write failed: ") + path);
  return GVX_OK;
}

static int file_map_read_file(gvx_map m, const char* path, int* side_matches)
{
  if (!m || !path) return fail(GVX_ERR_BAD_ARG, "gvx_map_read_file: null argument");
  std::lock_guard<std::recursive_mutex> g(m->mu);
  std::ifstream in(path, std::ios::in | std::ios::binary);
  if (!in.is_open()) return fail(GVX_ERR_BAD_ARG, std::string("cannot read ") + path);
  int32_t type = -1;
  float side = 0.f;
  uint32_t dims[3] = {0, 0, 0};
  in.read((char*)&type, 4);
  in.read((char*)&side, 4);
  in.read((char*)dims, 12);
  if (!in.good()) return fail(GVX_ERR_BAD_ARG, std::string("truncated map header: ") + path);
  if (type != m->type) return fail(GVX_ERR_BAD_ARG, "Voxelmap type does not match");
  if (dims[0] != m->dims[0] || dims[1] != m->dims[1] || dims[2] != m->dims[2])
    return fail(GVX_ERR_DIM_MISMATCH, "map dimensions in the file do not match");
  if (side_matches) *side_matches = side == m->side;  // a different side length only warns
  std::vector<char> buf(std::min<uint64_t>(m->bytes, kChunk));
  for (uint64_t o = 0; o < m->bytes; o += buf.size()) {
    const size_t len = (size_t)std::min<uint64_t>(buf.size(), m->bytes - o);
    in.read(buf.data(), (std::streamsize)len);
    if ((size_t)in.gcount() != len) return fail(GVX_ERR_BAD_ARG, std::string("truncated map data: ") + path);
    GVX_CHECK(copy_in((char*)m->data + o, buf.data(), len, GVX_MEM_HOST));
    GVX_CHECK(sync());
  }
  return GVX_OK;
}

static int file_list_write_file(gvx_list l, const char* path)
{
  if (!l || !path) return fail(GVX_ERR_BAD_ARG, "gvx_list_write_file: null argument");
  uint64_t n = 0;
  GVX_CHECK(gvx_list_size(l, &n));
  if (n > 0xFFFFFFFFull) return fail(GVX_ERR_BAD_ARG, "list too long for the 32-bit count of the format");
  uint32_t dims[3];
  float side;
  GVX_CHECK(gvx_list_info(l, dims, &side));
  std::vector<uint32_t> ids(n), coords(3 * n);
  std::vector<uint8_t> bits(32 * n);
  if (n) GVX_CHECK(gvx_list_read(l, ids.data(), coords.data(), bits.data(), GVX_MEM_HOST));
  std::ofstream out(path, std::ios::out | std::ios::binary | std::ios::trunc);
  if (!out.is_open()) return fail(GVX_ERR_BAD_ARG, std::string("cannot write ") + path);
  const int32_t type = GVX_MT_BITVECTOR_VOXELLIST;
  const uint32_t n32 = (uint32_t)n;
  out.write((const char*)&type, 4);
  out.write((const char*)dims, 12);
  out.write((const char*)&side, 4);
  out.write((const char*)&n32, 4);
  out.write((const char*)ids.data(), (std::streamsize)(4 * n));
  out.write((const char*)coords.data(), (std::streamsize)(12 * n));
  out.write((const char*)bits.data(), (std::streamsize)(32 * n));
  out.close();
  if (!out) return fail(GVX_ERR_BAD_ARG, std::string("write failed: ") + path);
  return GVX_OK;
}

static int file_list_read_file(gvx_list l, const char* path)
{
  if (!l || !path) return fail(GVX_ERR_BAD_ARG, "gvx_list_read_file: null argument");
  uint32_t dims[3];
  float side;
  GVX_CHECK(gvx_list_info(l, dims, &side));
  std::ifstream in(path, std::ios::in | std::ios::binary);
  if (!in.is_open()) return fail(GVX_ERR_BAD_ARG, std::string("cannot read ") + path);
  int32_t type = -1;
  uint32_t fd[3] = {0, 0, 0}, n = 0;
  float fs = 0.f;
  in.read((char*)&type, 4);
  if (!in.good() || type != GVX_MT_BITVECTOR_VOXELLIST) return fail(GVX_ERR_BAD_ARG, "list type does not match");
  in.read((char*)fd, 12);
  if (!in.good() || fd[0] != dims[0] || fd[1] != dims[1] || fd[2] != dims[2])
    return fail(GVX_ERR_DIM_MISMATCH, "reference map dimensions in the file do not match");
  in.read((char*)&fs, 4);
  if (!in.good() || fs != side) return fail(GVX_ERR_BAD_ARG, "voxel side length in the file does not match");
  in.read((char*)&n, 4);
  if (!in.good()) return fail(GVX_ERR_BAD_ARG, std::string("truncated list header: ") + path);
  if (48 * (uint64_t)n > bytes_left(in)) return fail(GVX_ERR_BAD_ARG, std::string("truncated list data: ") + path);
  std::vector<uint32_t> ids(n), coords(3 * (uint64_t)n);
  std::vector<uint8_t> bits(32 * (uint64_t)n);
  in.read((char*)ids.data(), (std::streamsize)(4 * (uint64_t)n));
  in.read((char*)coords.data(), (std::streamsize)(12 * (uint64_t)n));
  in.read((char*)bits.data(), (std::streamsize)(32 * (uint64_t)n));
  if (!in.good() && !(n == 0 && in.eof())) return fail(GVX_ERR_BAD_ARG, std::string("truncated list data: ") + path);
  return list_assign_host(l, n, ids.data(), coords.data(), bits.data());
}

}  // extern "C"

// exceptions (host allocation failures, stream errors) stop at the C boundary
template <class F>
static int no_throw(F f, const char* what)
{
  try {
    return f();
  } catch (const std::bad_alloc&) {
    return fail(GVX_ERR_OOM, std::string(what) + ": out of host memory");
  } catch (const std::exception& e) {
    return fail(GVX_ERR_BAD_ARG, std::string(what) + ": " + e.what());
  }
}

extern "C" {

int gvx_map_write_file(gvx_map m, const char* path)
{
  return no_throw([&] { return file_map_write_file(m, path); }, "gvx_map_write_file");
}

int gvx_map_read_file(gvx_map m, const char* path, int* side_matches)
{
  return no_throw([&] { return file_map_read_file(m, path, side_matches); }, "gvx_map_read_file");
}

int gvx_list_write_file(gvx_list l, const char* path)
{
  return no_throw([&] { return file_list_write_file(l, path); }, "gvx_list_write_file");
}

int gvx_list_read_file(gvx_list l, const char* path)
{
  return no_throw([&] { return file_list_read_file(l, path); }, "gvx_list_read_file");
}

}  // extern "C"
