// gvx_runtime.hip -- device selection, stream, staging, errors, timing, fills and scans.
#include <map>

#include "gvx_internal.hpp"

namespace gvx {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }
int fail(int code, const std::string& msg)
{
  g_err = msg;
  return code;
}

Runtime& rt()
{
  static Runtime r;
  return r;
}

int ensure_init()
{
  Runtime& r = rt();
  if (r.device >= 0) return GVX_OK;
  return gvx_init(0);
}

int dalloc(void** p, size_t bytes)
{
  *p = nullptr;
  if (bytes == 0) bytes = 4;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(GVX_ERR_OOM, std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
  }
  return GVX_OK;
}

void dfree(void* p)
{
  if (p) (void)hipFree(p);
}

int sync()
{
  GVX_HIP(hipStreamSynchronize(rt().stream));
  GVX_HIP(hipGetLastError());
  return GVX_OK;
}

int stage_reserve(size_t bytes, void** dev_out)
{
  Runtime& r = rt();
  if (bytes > r.stage_bytes) {
    GVX_CHECK(sync());
    dfree(r.stage);
    r.stage = nullptr;
    r.stage_bytes = 0;
    size_t want = bytes < (1u << 20) ? (1u << 20) : bytes + bytes / 4;
    GVX_CHECK(dalloc(&r.stage, want));
    r.stage_bytes = want;
  }
  *dev_out = r.stage;
  return GVX_OK;
}

int stage_input(const void* src, size_t bytes, int mem, const void** dev_out)
{
  if (mem == GVX_MEM_DEVICE || bytes == 0) {
    *dev_out = src;
    return GVX_OK;
  }
  void* d = nullptr;
  GVX_CHECK(stage_reserve(bytes, &d));
  GVX_HIP(hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, rt().stream));
  *dev_out = d;
  return GVX_OK;
}

int copy_out(void* dst, const void* dev_src, size_t bytes, int mem)
{
  if (bytes == 0) return GVX_OK;
  GVX_HIP(hipMemcpyAsync(dst, dev_src, bytes, mem == GVX_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                         rt().stream));
  return GVX_OK;
}

int copy_in(void* dev_dst, const void* src, size_t bytes, int mem)
{
  if (bytes == 0) return GVX_OK;
  GVX_HIP(hipMemcpyAsync(dev_dst, src, bytes, mem == GVX_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                         rt().stream));
  return GVX_OK;
}

// ---------------------------------------------------------------- profiling
struct ProfStore {
  std::mutex mu;
  std::map<std::string, std::vector<std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::map<std::string, std::pair<double, uint64_t>> done;
};
static ProfStore& prof()
{
  static ProfStore p;
  return p;
}

ProfScope::ProfScope(const char* n) : name(n)
{
  if (!n || !rt().profiling) return;  // no name: untimed
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
    a = b = nullptr;
    return;
  }
  (void)hipEventRecord(a, rt().stream);
}

ProfScope::~ProfScope()
{
  if (!a) return;
  (void)hipEventRecord(b, rt().stream);
  std::lock_guard<std::mutex> g(prof().mu);
  prof().pending[name].push_back({a, b});
}

static void prof_resolve()
{
  ProfStore& p = prof();
  std::lock_guard<std::mutex> g(p.mu);
  for (auto& kv : p.pending) {
    for (auto& ev : kv.second) {
      (void)hipEventSynchronize(ev.second);
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, ev.first, ev.second) == hipSuccess) {
        auto& d = p.done[kv.first];
        d.first += ms;
        d.second += 1;
      }
      (void)hipEventDestroy(ev.first);
      (void)hipEventDestroy(ev.second);
    }
    kv.second.clear();
  }
}

// ---------------------------------------------------------------- fills
__global__ void k_fill_u32(uint32_t* __restrict__ dst, uint32_t v, uint64_t n)
{
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t n4 = n / 4;
  uint4 vv = make_uint4(v, v, v, v);
  for (uint64_t k = i; k < n4; k += stride) reinterpret_cast<uint4*>(dst)[k] = vv;
  for (uint64_t k = n4 * 4 + i; k < n; k += stride) dst[k] = v;
}

__global__ void k_fill_bytes_tail(uint8_t* __restrict__ dst, uint8_t v, uint64_t n)
{
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = v;
}

int fill_u32(void* dst, uint32_t value, uint64_t n_words)
{
  if (n_words == 0) return GVX_OK;
  if (reinterpret_cast<uintptr_t>(dst) % 16 != 0) return fail(GVX_ERR_BAD_ARG, "fill_u32: misaligned");
  unsigned g = grid_for(n_words / 4 + 1, 256, 8192);
  hipLaunchKernelGGL(k_fill_u32, dim3(g), dim3(256), 0, rt().stream, (uint32_t*)dst, value, n_words);
  GVX_HIP(hipGetLastError());
  return GVX_OK;
}

int fill_bytes(void* dst, uint8_t value, uint64_t n_bytes)
{
  uint32_t v = value * 0x01010101u;
  uint64_t words = n_bytes / 4;
  if (words) GVX_CHECK(fill_u32(dst, v, words));
  uint64_t tail = n_bytes - words * 4;
  if (tail) {
    hipLaunchKernelGGL(k_fill_bytes_tail, dim3(1), dim3(64), 0, rt().stream, (uint8_t*)dst + words * 4, value, tail);
    GVX_HIP(hipGetLastError());
  }
  return GVX_OK;
}

// ---------------------------------------------------------------- exclusive scan (uint32)
// Three-level scheme: each 256-thread block scans 2048 elements (8 per thread),
// block totals are scanned recursively, then added back.
constexpr int SCAN_BLOCK = 256, SCAN_ITEMS = 8, SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* lds, uint32_t& total)
{
  // wave-level inclusive scan
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) lds[wid] = x;
  __syncthreads();
  uint32_t wbase = 0, tot = 0;
  for (int w = 0; w < SCAN_BLOCK / 64; ++w) {
    uint32_t s = lds[w];
    if (w < wid) wbase += s;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return wbase + x - v;
}

__global__ void k_scan_tiles(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint64_t n,
                             uint32_t* __restrict__ block_sums)
{
  __shared__ uint32_t lds[SCAN_BLOCK / 64];
  uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    uint64_t i = base + k;
    v[k] = i < n ? in[i] : 0u;
    s += v[k];
  }
  uint32_t tot;
  uint32_t pre = block_excl_scan(s, lds, tot);
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    uint64_t i = base + k;
    if (i < n) out[i] = pre;
    pre += v[k];
  }
  if (threadIdx.x == 0 && block_sums) block_sums[blockIdx.x] = tot;
}

__global__ void k_scan_add(uint32_t* __restrict__ out, uint64_t n, const uint32_t* __restrict__ block_offs)
{
  uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
  uint32_t add = block_offs[blockIdx.x];
  for (int k = threadIdx.x; k < SCAN_TILE; k += SCAN_BLOCK) {
    uint64_t i = base + k;
    if (i < n) out[i] += add;
  }
}

__global__ void k_scan_total(const uint32_t* __restrict__ in, const uint32_t* __restrict__ out, uint64_t n,
                             uint32_t* __restrict__ total)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) *total = n ? out[n - 1] + in[n - 1] : 0u;
}

struct ScanScratch {
  std::vector<void*> lvl;
  std::vector<size_t> cap;
};
static ScanScratch& scan_scratch()
{
  static ScanScratch s;
  return s;
}

static int scan_level(const uint32_t* in, uint32_t* out, uint64_t n, int depth)
{
  uint64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (tiles <= 1) {
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(SCAN_BLOCK), 0, rt().stream, in, out, n, (uint32_t*)nullptr);
    GVX_HIP(hipGetLastError());
    return GVX_OK;
  }
  ScanScratch& ss = scan_scratch();
  if ((int)ss.lvl.size() <= depth) {
    ss.lvl.resize(depth + 1, nullptr);
    ss.cap.resize(depth + 1, 0);
  }
  size_t need = 2 * tiles * sizeof(uint32_t);
  if (ss.cap[depth] < need) {
    GVX_CHECK(sync());
    dfree(ss.lvl[depth]);
    GVX_CHECK(dalloc(&ss.lvl[depth], need));
    ss.cap[depth] = need;
  }
  uint32_t* sums = (uint32_t*)ss.lvl[depth];
  uint32_t* offs = sums + tiles;
  hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)tiles), dim3(SCAN_BLOCK), 0, rt().stream, in, out, n, sums);
  GVX_HIP(hipGetLastError());
  GVX_CHECK(scan_level(sums, offs, tiles, depth + 1));
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)tiles), dim3(SCAN_BLOCK), 0, rt().stream, out, n, offs);
  GVX_HIP(hipGetLastError());
  return GVX_OK;
}

int scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* total_dev)
{
  if (n > 0) GVX_CHECK(scan_level(in, out, n, 0));
  if (total_dev) {
    hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(64), 0, rt().stream, in, out, n, total_dev);
    GVX_HIP(hipGetLastError());
  }
  return GVX_OK;
}

// ---------------------------------------------------------------- thresholds (host, unfused)
#pragma clang fp contract(off)
int8_t float_to_probability(float v)
{
  float t = v < 1.0f ? v : 1.0f;
  t = t > 0.0f ? t : 0.0f;
  volatile float prod = t * ((float)PROB_MAX - (float)PROB_MIN);
  float r = prod + (float)PROB_MIN;
  return (int8_t)r;
}

int8_t sv_float_to_probability(float v)
{
  volatile float prod = v * ((float)PROB_MAX - (float)PROB_MIN);
  float t = prod + (float)PROB_MIN;
  if (t >= 127.0f) return 127;
  if (t <= -128.0f) return -128;
  return (int8_t)t;
}

}  // namespace gvx

using namespace gvx;

extern "C" {

int gvx_init(int device)
{
  Runtime& r = rt();
  std::lock_guard<std::recursive_mutex> g(r.mu);
  if (r.device >= 0) {
    if (r.device == device) return GVX_OK;
    return fail(GVX_ERR_BAD_ARG, "gvx_init: already initialised on another device");
  }
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    (void)hipGetLastError();
    return fail(GVX_ERR_NO_DEVICE, "gvx_init: no HIP device");
  }
  if (device < 0 || device >= count) return fail(GVX_ERR_BAD_ARG, "gvx_init: bad device index");
  GVX_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  GVX_HIP(hipGetDeviceProperties(&prop, device));
  r.num_cus = prop.multiProcessorCount;
  GVX_HIP(hipStreamCreateWithFlags(&r.stream, hipStreamNonBlocking));
  GVX_CHECK(dalloc(&r.dev_small, 4096));
  GVX_HIP(hipHostMalloc(&r.host_small, 4096, hipHostMallocDefault));
  r.device = device;
  // device address of the EDT's fixed workspace; a failure is reported by the
  // one entry point that needs it, not here
  if (pba_noalloc_setup(device) != GVX_OK) set_error("");
  return GVX_OK;
}

const char* gvx_last_error(void) { return g_err.c_str(); }

int gvx_get_stream(void** stream)
{
  GVX_CHECK(ensure_init());
  if (!stream) return fail(GVX_ERR_BAD_ARG, "null");
  *stream = (void*)rt().stream;
  return GVX_OK;
}

int gvx_profile_enable(int enable)
{
  GVX_CHECK(ensure_init());
  rt().profiling = enable != 0;
  return GVX_OK;
}

int gvx_profile_get(const char* kernel, double* avg_ms, uint64_t* launches)
{
  if (!kernel) return fail(GVX_ERR_BAD_ARG, "null kernel name");
  prof_resolve();
  std::lock_guard<std::mutex> g(prof().mu);
  auto it = prof().done.find(kernel);
  if (it == prof().done.end() || it->second.second == 0) {
    if (avg_ms) *avg_ms = 0.0;
    if (launches) *launches = 0;
    return GVX_OK;
  }
  if (avg_ms) *avg_ms = it->second.first / (double)it->second.second;
  if (launches) *launches = it->second.second;
  return GVX_OK;
}

int gvx_profile_reset(void)
{
  prof_resolve();
  std::lock_guard<std::mutex> g(prof().mu);
  prof().done.clear();
  return GVX_OK;
}

}  // extern "C"
