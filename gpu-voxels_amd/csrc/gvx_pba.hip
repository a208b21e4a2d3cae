// gvx_pba.hip -- DistanceVoxelMap: exact 3D EDT (PBA semantics), distance read-outs,
// mergeOccupied and the brute-force exactDistances3D.
//
// The EDT runs as three axis passes with no transposes of the map:
//   z pass  one thread per (x, y) column; the column's occupancy is held as a
//           register bitmask (all loads issued before any use), so each voxel is
//           read once and the nearest obstacle z (ties to the lower z) is written
//           as uint16.
//   y pass  on dense fields a windowed search (k_pba_y_win): every voxel takes the
//           lowest key cost + k^2 within a radius that grows until the whole wave
//           has resolved, and stores its site relative to itself in 2 bytes.  Z
//           slices it cannot resolve (far voxels, sparse fields) go to the
//           proximate-point stack kernel (k_pba_y: PBA phase 2 and phase 3 with
//           the stack's top entries in an LDS ring per line, spilled to a per-line
//           global array only when deeper), which writes packed DistanceVoxels.
//   x pass  the same windowed search along x (k_pba_x_win), reading each slice's
//           y output in its format; rows it cannot resolve go to the stack kernel
//           (k_pba_x, 64 x 16 LDS tiles turning the per-lane walk into row loads,
//           final stacks to a global array) and the colouring kernel
//           (k_pba_x_colour: each position's owner as a prefix maximum of the
//           entries' first positions, whole row segments written).
// All bisector tests are exact integer comparisons.
#include "gvx_internal.hpp"

namespace gvx {

constexpr uint16_t ZNONE = 0xFFFF;

// ---------------------------------------------------------------- z pass
// Block = 64 columns (lanes = consecutive (x, y) columns) x ZP waves; wave w takes
// the z range [128 w, 128 w + 128) of every column: its occupancy as 4 register
// words (all 128 loads issued before any is used).  The words of all waves meet in
// LDS, so each wave also knows the nearest obstacle below and above its range,
// and writes the nearest obstacle z (ties to the lower z) of its range as uint16.
// FULL: dz == 32 * W, every load unconditional so the compiler keeps many in flight.
// With dx % 64 == 0, colm[(y / 32) * (dx / 64) + x / 64] gets bit y % 32 when the
// 64 columns (x, y) of that row segment were stored (one of them holds an obstacle);
// the other segments are not stored.  Other widths store every column.
// tflags[x / 64] gets bit y / tile_rows for every column (x, y) holding an
// obstacle: the y pass skips (x range, row tile) blocks without one.
// xcnt[x / 64] counts the obstacles of each 64-column x range (the y pass's density test).
constexpr int ZWP = 4;  // occupancy words (of 32 z) per wave
template <int W, bool FULL>
__global__ void __launch_bounds__(64 * (W / ZWP)) k_pba_z(const uint32_t* __restrict__ map,
                                                          uint16_t* __restrict__ zs, uint32_t* __restrict__ tflags,
                                                          uint32_t* __restrict__ colm, uint32_t* __restrict__ xcnt,
                                                          uint32_t tile_rows, uint32_t dx, uint32_t dy, uint32_t dz,
                                                          uint32_t uninit)
{
  static_assert(W % ZWP == 0, "whole parts");
  __shared__ uint32_t sb[W][64];  // occupancy word w of the 64 columns
  const uint32_t lane = threadIdx.x & 63, part = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const uint64_t col = (uint64_t)blockIdx.x * 64 + lane;
  const bool cin = col < (uint64_t)dx * dy;
  const uint64_t layer = (uint64_t)dx * dy;
  const uint32_t* __restrict__ p = map + (cin ? col : 0);
  uint32_t bits[ZWP];
#pragma unroll
  for (int i = 0; i < ZWP; ++i) {
    const int w = (int)part * ZWP + i;
    uint32_t v[32];
    if (FULL) {
#pragma unroll
      for (int b = 0; b < 32; ++b) v[b] = p[(uint64_t)(w * 32 + b) * layer];
    } else {
#pragma unroll
      for (int b = 0; b < 32; ++b) {
        const uint32_t z = (uint32_t)(w * 32 + b);
        v[b] = z < dz ? p[(uint64_t)z * layer] : uninit;
      }
    }
    uint32_t word = 0;
#pragma unroll
    for (int b = 0; b < 32; ++b) word |= (v[b] != uninit ? 1u : 0u) << b;
    bits[i] = cin ? word : 0u;
    sb[w][lane] = bits[i];
  }
  __syncthreads();
  // nearest obstacle below this range (last set bit of the lower words) and above it
  // (first set bit of the higher words); 0xFFFF = none
  uint32_t last = 0xFFFFu, after = 0xFFFFu;
  for (int w = (int)part * ZWP - 1; w >= 0 && last == 0xFFFFu; --w) {
    const uint32_t m = sb[w][lane];
    if (m) last = (uint32_t)(w * 32 + 31 - __builtin_clz(m));
  }
  for (int w = ((int)part + 1) * ZWP; w < W && after == 0xFFFFu; ++w) {
    const uint32_t m = sb[w][lane];
    if (m) after = (uint32_t)(w * 32 + __builtin_ctz(m));
  }
  uint32_t any = last != 0xFFFFu || after != 0xFFFFu ? 1u : 0u;
#pragma unroll
  for (int i = 0; i < ZWP; ++i) any |= bits[i];
  // with dx % 64 == 0 a block is 64 columns of one row: when none of them holds an
  // obstacle nothing is stored (its segment's bit in colm stays clear and the y
  // pass never reads it); with one, all 64 are, so stores stay whole rows
  const bool seg = dx % 64 == 0;
  const bool wave_any = wave_ballot(any != 0u) != 0ull;
  if (cin && (!seg || wave_any)) {
    uint16_t* __restrict__ o = zs + col;
#pragma unroll
    for (int i = 0; i < ZWP; ++i) {
      const int w = (int)part * ZWP + i;
      // first set position in the words after this one (in this range, else above it)
      uint32_t nxt = after;
#pragma unroll
      for (int v = ZWP - 1; v > i; --v) nxt = bits[v] ? (uint32_t)(((int)part * ZWP + v) * 32 + __builtin_ctz(bits[v])) : nxt;
#pragma unroll
      for (int b = 0; b < 32; ++b) {
        const uint32_t z = (uint32_t)(w * 32 + b);
        if (FULL || z < dz) {
          if ((bits[i] >> b) & 1u) last = z;
          const uint32_t m = bits[i] & (0xFFFFFFFFu << b);
          const uint32_t hi = m ? (uint32_t)(w * 32 + __builtin_ctz(m)) : nxt;
          uint32_t pick;
          if (last == 0xFFFFu) pick = hi;
          else if (hi == 0xFFFFu) pick = last;
          else pick = (z - last <= hi - z) ? last : hi;
          o[(uint64_t)z * layer] = (uint16_t)pick;
        }
      }
    }
  }
  if (part != 0 || !cin) return;
  uint32_t nob = 0;
  for (int w = 0; w < W; ++w) nob += (uint32_t)__popc(sb[w][lane]);
  const uint32_t c32 = (uint32_t)col, cx = c32 % dx, cy = c32 / dx;  // col < 2^20
  const uint32_t fw = cx / 64, fb = 1u << (cy / tile_rows);
  if (seg) {  // the block's 64 columns share one flag bit and one segment bit: one atomic each
    for (int o = 32; o > 0; o >>= 1) nob += (uint32_t)__shfl_xor((int)nob, o, 64);
    if (lane == 0 && wave_any) {
      atomicOr(tflags + fw, fb);
      atomicOr(colm + (uint64_t)(cy >> 5) * (dx / 64) + cx / 64, 1u << (cy & 31));
      atomicAdd(xcnt + fw, nob);
    }
  } else if (any) {
    atomicOr(tflags + fw, fb);
    atomicAdd(xcnt + fw, nob);
  }
}

// columns longer than 1024: two sweeps through the uint16 buffer
__global__ void __launch_bounds__(256) k_pba_z_long(const uint32_t* __restrict__ map, uint16_t* __restrict__ zs,
                                                    uint32_t dx, uint32_t dy, uint32_t dz, uint32_t uninit)
{
  const uint64_t col = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= (uint64_t)dx * dy) return;
  const uint64_t layer = (uint64_t)dx * dy;
  uint32_t last = 0xFFFFu;
  for (uint32_t z = 0; z < dz; ++z) {
    if (map[z * layer + col] != uninit) last = z;
    zs[z * layer + col] = (uint16_t)last;
  }
  uint32_t next = 0xFFFFu;
  for (int64_t zz = (int64_t)dz - 1; zz >= 0; --zz) {
    const uint32_t z = (uint32_t)zz;
    if (map[z * layer + col] != uninit) next = z;
    const uint32_t lo = zs[z * layer + col];
    uint32_t pick;
    if (lo == 0xFFFFu) pick = next;
    else if (next == 0xFFFFu) pick = lo;
    else pick = (z - lo <= next - z) ? lo : next;
    zs[z * layer + col] = (uint16_t)pick;
  }
}

// ---------------------------------------------------------------- y and x passes
// One wave per block handles 64 lines (lanes = lines) and runs PBA's phase 2
// (proximate-point stack) and phase 3 (colouring, from the top of the line down).
// The stack of a line lives in an LDS ring of its RK topmost entries; entries
// pushed deeper spill to a per-line global array (stk[line * n + depth]) and come
// back RK - 1 at a time in one batch of independent loads when a pop reaches
// below the ring.  Short stacks therefore never touch global memory.  The input
// is staged through LDS tiles so global loads and stores are row-contiguous, and
// each lane walks only the positions of a tile that hold a site (a bit mask), so
// empty stretches of a line cost nothing.
//
// Bisector tests are exact integer arithmetic.  For sites a < b on the line
// (positions pa, pb; squared off-line costs ca, cb), a is at least as close as b
// at position p exactly when p * D <= N with N = (pb^2 - pa^2) + cb - ca and
// D = 2 (pb - pa) > 0.  Positions are < 1024 and costs < 2^21, so N < 2^23 and
// p * D < 2^21 fit int32; phase 2's comparison N_ab * D_bc >= N_bc * D_ab (pop b,
// also on equality) has products < 2^35, exact in double precision.
#ifndef GVX_PBA_TYR
#define GVX_PBA_TYR 32
#endif
#ifndef GVX_PBA_TX
#define GVX_PBA_TX 16
#endif
#ifndef GVX_PBA_RK
#define GVX_PBA_RK 16
#endif
constexpr int TB = 64;            // lines per wave
constexpr int TYR = GVX_PBA_TYR;  // y pass: positions per input tile
constexpr int TX = GVX_PBA_TX;    // x pass: positions per tile (4 TX-byte row segments)
constexpr int RK = GVX_PBA_RK;    // stack entries per line held in LDS
static_assert(RK >= 4 && (RK & (RK - 1)) == 0 && TX >= 4 && 64 % TX == 0 && TYR <= 32, "tile shape");

__device__ __forceinline__ int32_t sq24(int32_t v) { return __mul24(v, v); }

// raw buffer loads over one z slice of a map or of the z-pass output (below
// 4 GiB): a per-lane byte offset plus a wave-uniform byte offset, so the tile
// loads need no per-load vector address arithmetic
__device__ __forceinline__ __amdgpu_buffer_rsrc_t slice_rsrc(const void* base, uint32_t bytes)
{
  const uint64_t b = reinterpret_cast<uint64_t>(base);  // the same for the whole wave: keep it scalar
  const uint64_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)b);
  const uint64_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(b >> 32));
  return __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void*>(hi << 32 | lo), (short)0,
                                           __builtin_amdgcn_readfirstlane((int)bytes), 0x00020000);
}

// bisector of sites a < b: a is at least as close as b at p  <=>  p * D <= N
struct Bis {
  int32_t N, D;
  __device__ __forceinline__ void set(int32_t pa, int32_t ca, int32_t pb, int32_t cb)
  {
    D = 2 * (pb - pa);
    N = __mul24(pb - pa, pb + pa) + (cb - ca);
  }
  __device__ __forceinline__ bool a_wins(int32_t p) const { return __mul24(p, D) <= N; }
};
// b (the middle of a < b < c) owns no position: bisector(a, b) >= bisector(b, c)
__device__ __forceinline__ bool middle_empty(const Bis& ab, const Bis& bc)
{
  return (double)ab.N * (double)bc.D >= (double)bc.N * (double)ab.D;
}

// Per-lane stack of 32-bit entries: ring[depth % RK][lane] holds depths [lo, n);
// depths below lo are in g[0 .. lo).  The caller also keeps the top three
// entries in registers, so the only fetch is that of the new third entry after
// a pop.
struct LineStack {
  uint32_t (*ring)[TB];
  uint32_t* g;
  uint32_t lane, n, lo;

  __device__ __forceinline__ void push(uint32_t e)
  {
    if (n - lo == (uint32_t)RK) {  // ring full: the slot of depth n holds depth lo
      g[lo] = ring[lo % RK][lane];
      ++lo;
    }
    ring[n % RK][lane] = e;
    ++n;
  }
  __device__ __forceinline__ void pop() { --n; }  // lo <= n - 2 before a pop (see at3)
  // entry at depth d = n - 3 (any d <= n - 3 is harmless when n < 3: then
  // lo == 0 and nothing is reloaded): the caller keeps the top two entries in
  // registers and this third one too, fetching it right after a pop so the LDS
  // read is in flight while the next test runs.  Below the ring, depths
  // d+3-RK .. d come back in one batch of independent loads; the slots of
  // depths n - 2 and n - 1 are left alone (the ring always holds them: refills
  // happen as soon as d < lo, which keeps lo <= max(n - 2, 0)).
  __device__ __forceinline__ uint32_t at3(uint32_t d)
  {
    if (d < lo) {
      uint32_t v[RK - 2];  // all loads first, so they are in flight together
#pragma unroll
      for (int j = 0; j < RK - 2; ++j) v[j] = g[(uint32_t)j <= d ? d - (uint32_t)j : d];
#pragma unroll
      for (int j = 0; j < RK - 2; ++j) ring[((uint32_t)j <= d ? d - (uint32_t)j : d) % RK][lane] = v[j];
      lo = d + 3 >= (uint32_t)RK ? d + 3 - RK : 0;
    }
    return ring[d % RK][lane];
  }
};


// y pass: lines (x, z), lanes = consecutive x.  Stack entry = y | z_site << 10.
// Every lane touches only its own LDS columns, so no barriers are needed.
// Output: the packed DistanceVoxel (x, y_site, z_site) written into the map.
// Row tiles whose 64 x 32 block of columns holds no obstacle (tflags, from the z
// pass) are neither loaded nor walked: they hold no site.  xflags[z] gets bit
// x / TX for every line (x, z) left with a site, for the x pass.
__global__ void __launch_bounds__(TB) k_pba_y(const uint16_t* __restrict__ zs, uint32_t* __restrict__ out,
                                              uint32_t* __restrict__ stk, const uint32_t* __restrict__ tflags,
                                              const uint32_t* __restrict__ colm,
                                              unsigned long long* __restrict__ xflags,
                                              uint32_t* __restrict__ yfs, uint32_t dx, uint32_t dy,
                                              uint32_t dz, uint32_t uninit)
{
  static_assert(TYR == 32, "one column-mask word per row tile");
  __shared__ uint16_t ztile[TYR][TB];
  __shared__ uint32_t ring[RK][TB];
  const uint32_t ngroups = (dx + TB - 1) / TB * dz;
  if (blockIdx.x < dz) {
    // the format of slice z's y output for the x pass (yfs[dz + ngroups + z]): 0 =
    // relative sites, 1 = packed (the slice, or each of its groups, is redone here),
    // 2 = both
    const uint32_t zz = blockIdx.x, nxr = (dx + TB - 1) / TB, l = threadIdx.x;
    const uint64_t fm = wave_ballot(l < nxr && yfs[dz + l * dz + zz] != 0u);
    const uint64_t allc = nxr >= 64 ? ~0ull : (1ull << nxr) - 1ull;
    if (l == 0) yfs[dz + ngroups + zz] = yfs[zz] != 0u || fm == allc ? 1u : fm == 0ull ? 0u : 2u;
  }
  // the windowed kernel resolved this group: neither its slice (yfs[z]) nor the group
  // itself (yfs[dz + group]) was left to this kernel
  if (yfs[blockIdx.x % dz] == 0u && yfs[dz + blockIdx.x] == 0u) return;
  // consecutive blocks take consecutive z with the same x range: workgroups are
  // dealt round-robin over the XCDs, so every XCD gets every x range (obstacles
  // clustered in a few x ranges would otherwise load a few XCDs only)
  const uint32_t z = blockIdx.x % dz, x0 = (blockIdx.x / dz) * TB;
  const uint32_t lane = threadIdx.x, x = x0 + lane;
  if (x >= dx) return;
  const int32_t zi = (int32_t)z;
  LineStack st{ring, stk + ((uint64_t)z * dx + x) * dy, lane, 0, 0};
  // top (ty, tz, tc), second (sy, sz, sc) and third (entry e3) of the stack;
  // bst = bisector(second, top)
  int32_t ty = 0, tz = 0, tc = 0, sy = 0, sz = 0, sc = 0;
  uint32_t e3 = 0;
  Bis bst{0, 1};
  // pop the top: second and third move up, the new third is fetched (when fewer
  // than two entries remain the moved values are never read)
  auto pop_y = [&]() {
    st.pop();
    ty = sy;
    tz = sz;
    tc = sc;
    sy = (int32_t)(e3 & 0x3FFu);
    sz = (int32_t)(e3 >> 10);
    sc = sq24(sz - zi);
    bst.set(sy, sc, ty, tc);
    e3 = st.at3(st.n >= 3 ? st.n - 3 : 0);
  };
  // software pipeline: the next tile's values are loaded into registers (rows
  // clamped into the line) while the current tile is processed
  const __amdgpu_buffer_rsrc_t zr = slice_rsrc(zs + (uint64_t)z * dx * dy, dx * dy * 2);
  auto zload = [&](uint32_t yr) -> uint32_t {  // row yr of this lane's line, yr < dy
    return __builtin_amdgcn_raw_buffer_load_b16(zr, (int)(x * 2), (int)(yr * dx * 2), 0);
  };
  // row tiles with an obstacle column in this x range, in increasing y
  const uint32_t ntiles = (dy + TYR - 1) / TYR;
  uint32_t todo = tflags[x0 / 64] & (ntiles >= 32 ? 0xFFFFFFFFu : (1u << ntiles) - 1u);
  uint32_t pre[TYR], pre_cm = 0;
  auto load_rows = [&](uint32_t t) {  // past the end of the line: harmless reloads of the last row
    // rows of this tile whose 64-column segment was stored (TYR == 32; all for other widths)
    pre_cm = dx % 64 == 0 ? colm[(uint64_t)t * (dx / 64) + x0 / 64] : 0xFFFFFFFFu;
#pragma unroll
    for (int r = 0; r < TYR; ++r) {
      const uint32_t yr = t * TYR + r;
      pre[r] = zload(yr < dy ? yr : dy - 1);
    }
  };
  if (todo) load_rows((uint32_t)__builtin_ctz(todo));
  while (todo) {
    const uint32_t t = (uint32_t)__builtin_ctz(todo);
    todo &= todo - 1u;
    const uint32_t y0 = t * TYR;
    const uint32_t rows = dy - y0 < (uint32_t)TYR ? dy - y0 : (uint32_t)TYR;
    const uint32_t cm = pre_cm;
    uint32_t m = 0;  // positions of this tile that hold a site
#pragma unroll
    for (int r = 0; r < TYR; ++r) {
      ztile[r][lane] = (uint16_t)pre[r];
      m |= (pre[r] != ZNONE ? 1u : 0u) << r;
    }
    m &= cm;  // a segment without an obstacle was not stored (its values are stale)
    if (rows < (uint32_t)TYR) m &= (1u << rows) - 1u;
    load_rows(todo ? (uint32_t)__builtin_ctz(todo) : t);  // the next tile (the last one reloads itself)
    // the next site's value is read from LDS while the current one is processed
    uint32_t r = m ? (uint32_t)__builtin_ctz(m) : 0u;
    int32_t s = ztile[r][lane];
    while (m) {
      m &= m - 1u;
      const uint32_t rn = m ? (uint32_t)__builtin_ctz(m) : r;
      const int32_t sn = ztile[rn][lane];
      const int32_t y = (int32_t)(y0 + r), c = sq24(s - zi);
      if (st.n >= 1) {
        Bis btn;
        btn.set(ty, tc, y, c);
        while (st.n >= 2 && middle_empty(bst, btn)) {
          pop_y();
          btn.set(ty, tc, y, c);
        }
        bst = btn;
      }
      st.push((uint32_t)y | ((uint32_t)s << 10));
      e3 = (uint32_t)sy | ((uint32_t)sz << 10);
      sy = ty;
      sz = tz;
      sc = tc;
      ty = y;
      tz = s;
      tc = c;
      r = rn;
      s = sn;
    }
  }
  // x tiles (TX lines each) of this slice that keep a site
  const uint64_t kept = wave_ballot(st.n > 0);
  {
    unsigned long long f = 0;
#pragma unroll
    for (int k = 0; k < TB / TX; ++k)
      if ((kept >> (k * TX)) & ((1ull << TX) - 1ull)) f |= 1ull << (x0 / TX + k);
    if (lane == 0 && f) atomicOr(xflags + z, f);
  }
  // a line whose x tile kept no site is never read again: the x pass skips the tile
  // and the colouring writes every voxel of every row, so its rows are not stored
  if (((kept >> ((lane / TX) * TX)) & ((1ull << TX) - 1ull)) == 0) return;
  // output rows through a descriptor over this z slice: lane offset x, row offset uniform
  const __amdgpu_buffer_rsrc_t orr = slice_rsrc(out + (uint64_t)z * dx * dy, dx * dy * 4);
  if (st.n == 0) {
    for (uint32_t y = 0; y < dy; ++y) __builtin_amdgcn_raw_buffer_store_b32(uninit, orr, (int)(x * 4), (int)(y * dx * 4), 0);
    return;
  }
  for (int32_t yy = (int32_t)dy - 1; yy >= 0; --yy) {
    while (st.n >= 2 && bst.a_wins(yy)) {
      pop_y();
    }
    __builtin_amdgcn_raw_buffer_store_b32(x | ((uint32_t)ty << 10) | ((uint32_t)tz << 20), orr, (int)(x * 4),
                                          (int)((uint32_t)yy * dx * 4), 0);
  }
}

// ---------------------------------------------------------------- windowed passes
// The three passes give every voxel its nearest obstacle with ties broken to the
// lowest coordinate along the pass's axis (z flood: lower z; the proximate-point
// stack pops on equality and colours ties to the lower entry), so the result is
// the nearest obstacle with ties to the lexicographically smallest (x, y, z).
//
// On fields where every voxel has a site within a few voxels, a search of the
// window |q - p| <= k finds a position's winner directly: every site outside the
// window costs more than k^2, so once the window's minimum is below (k + 1)^2
// nothing farther can reach or tie it.  Each site carries one 32-bit key,
//   cost << SH | position << (SH - 10) | payload,
// so the window minimum of key + (k^2 << SH) is the lowest minimum cost and,
// among equal costs, the lowest position (the tie rule) in one v_min3 per step;
// the payload rides along (positions are unique).  The x pass uses SH = 11 (a
// 20-bit cost, no payload), the y pass SH = 20 (an 11-bit cost and the site's z
// as payload, so the winner's key is its site).  Costs saturate at the field's
// maximum: a winner costs less than (WIN_R + 1)^2 = 1089, so a saturated site
// never resolves a voxel.  Lanes take consecutive
// positions of a line held in LDS and grow the radius together, one ds_read2 and
// one min3 per step, until every lane has resolved (or the radius reaches
// WIN_R).  A group with a voxel that does not resolve within WIN_R (or whose
// block obviously holds far voxels: part of it has no obstacle) is flagged, and
// the proximate-point kernels redo the whole group (in the y pass: its whole z
// slice).  The output is the same either way.
constexpr int WIN_R = 32;                    // largest window radius
constexpr uint32_t KEY_NONE = 0x7FFFFFFFu;   // no site: at or above every real key, below 2^31
constexpr int WIN_CHECK = 4;                 // radius steps between the wave's all-resolved tests

// x-pass key (SH = 11): costs from 2^20 - 1 up saturate, so the key stays below 2^31
__device__ __forceinline__ uint32_t site_key(uint32_t cost, uint32_t pos)
{
  return ((cost < 0xFFFFFu ? cost : 0xFFFFFu) << 11) | (pos << 1);  // pos < 1024
}

// minimum over |q - p| <= k of kp[R + q - p] + ((q - p)^2 << SH), k growing until
// every active lane of the wave has resolved (its minimum cost below (k + 1)^2) or
// k = kmax (<= WIN_R); kp points at position p - WIN_R of a key array with kmax valid
// or KEY_NONE entries on each side.  ok: resolved.
template <int SH>
__device__ __forceinline__ uint32_t win_search(const uint32_t* kp, bool active, bool& ok, int kmax = WIN_R)
{
  // keys are below 2^31, so they compare the same as int32.  b = (minimum so far) -
  // (k^2 << SH): going from radius k - 1 to k subtracts (2k - 1) << SH and takes the
  // min with the two new keys, one subtract and one v_min3 per step (b stays above
  // -(WIN_R^2 << SH) >= -2^30).  A lane is resolved at radius k once its minimum is
  // below k^2 << SH, i.e. b < (2k - 1) << SH (it stays so as k grows); an inactive
  // lane starts resolved (key 0).
  const int32_t* ip = reinterpret_cast<const int32_t*>(kp);
  int32_t b = ip[WIN_R];  // kp is a valid address for every lane
  b = active ? b : 0;
  int k = 1;
#pragma unroll
  for (; k <= WIN_R; ++k) {
    // kmax: a multiple of WIN_CHECK
    if (k % WIN_CHECK == 1 && (k > kmax || __builtin_amdgcn_ballot_w64(b >= ((2 * k - 1) << SH)) == 0)) break;
    b = min(min(b - ((2 * k - 1) << SH), ip[WIN_R - k]), ip[WIN_R + k]);
  }
  // radius k - 1 was the last step; unsigned: with SH = 20 a saturated minimum plus k^2 passes 2^31
  const uint32_t best = (uint32_t)b + ((uint32_t)((k - 1) * (k - 1)) << SH);
  ok = best < ((uint32_t)((kmax + 1) * (kmax + 1)) << SH);
  return best;
}

// y pass, windowed: a block takes 64 lines x0 .. x0+63 at one z and two row blocks
// of 64 output rows, y0 .. y0+63 and y0+64 .. y0+127.  For each row block the z-pass
// rows y0 - H .. y0 + 63 + H of the 64 columns are loaded with coalesced row loads
// (the second row block's while the first is searched) and stored transposed (one
// LDS row of keys per line, key position = y, payload = the site's z); each wave
// searches 8 output rows with lanes = the 64 lines (lane l reads line l's keys, the
// odd line stride keeps them on distinct banks), so its winners, whose keys hold
// their sites, are whole rows and are stored directly.  The halo H follows the
// x range's obstacle density rho (counted by the z pass): H = 2.5 / sqrt(rho),
// rounded up to WIN_CHECK, at most WIN_R (28 at 1 %, where 8 of the 32768 row blocks
// of a uniform 512^3 field need more, 8 at 10 %, where 3 do); a row block that does
// not resolve is staged again with H = WIN_R, and only one that fails then sends its
// z slice to the stack kernel.  Below rho = 1/256 (or with an obstacle-free row tile)
// the x range's groups go to the stack kernel.
// Blocks are dealt to the XCDs so that the blocks of one (x range, z) group run one
// after another on one XCD, where the overlapping halo rows are still in its L2.
constexpr int YB = 64;                 // output rows per row block
constexpr int YT = YB + 2 * WIN_R;     // rows staged per block, at most
constexpr int YS = YT + 1;             // LDS words per staged line (odd: transposed writes spread over the banks)
constexpr int YW_WAVES = 8;
static_assert(32 % YW_WAVES == 0, "a wave's staged rows r = w + YW_WAVES i share word (YW_WAVES i) / 32");
constexpr int YW_RPS = YB / YW_WAVES;  // output rows searched per wave (lanes = the 64 lines)
// n / d and n % d for a wave-uniform divisor d >= 1 with m = 0xFFFFFFFF / d from the
// host: mulhi(n, m) is q or q - 1, one correction makes it exact (a handful of
// scalar instructions instead of a full division)
struct UDiv {
  uint32_t d, m;
  __device__ __forceinline__ uint32_t div(uint32_t n, uint32_t& r) const
  {
    uint32_t q = __umulhi(n, m);
    r = n - q * d;
    if (r >= d) {
      ++q;
      r -= d;
    }
    return q;
  }
};
static inline UDiv udiv(uint32_t d) { return UDiv{d, 0xFFFFFFFFu / d}; }

// The windowed y pass writes each voxel's site relative to the voxel, 2 bytes:
// (y_site - y + 64) | (z_site - z + 64) << 7.  A resolved window search finds its
// site within WIN_R rows and at a cost below (WIN_R + 1)^2, so both offsets lie in
// [-WIN_R, WIN_R].  The stack kernel redoes every z slice with a group the window
// left unresolved, and the groups of x ranges too sparse for the window, and writes
// packed DistanceVoxels to the 4-byte buffer; the x pass reads a slice's rows in its
// one format, or leaves rows whose slice has both to the x stack kernels.
static_assert(WIN_R < 64, "relative sites hold offsets in [-64, 63]");
// For a resolved window winner (key low bits y_site << 10 | z_site) neither field
// wraps, so the relative site is a sum: c = yrel_base(y, z)
__device__ __forceinline__ uint32_t yrel_base(uint32_t y, uint32_t z) { return (64u - y) + ((64u - z) << 7); }
__device__ __forceinline__ uint32_t yrel_win(uint32_t key, uint32_t c)
{
  return ((key >> 10) & 1023u) + ((key & 1023u) << 7) + c;
}
// the packed DistanceVoxel (x, y_site, z_site) of a relative site at (x, y, z)
__device__ __forceinline__ uint32_t yrel_site(uint32_t rel, uint32_t x, uint32_t y, uint32_t z)
{
  return x | ((y + (rel & 127u) - 64u) << 10) | ((z + (rel >> 7) - 64u) << 20);
}

// Row block stages of the y window (halo H, a multiple of WIN_CHECK): wave w stages
// rows y = y0 - H + w + YW_WAVES i.  ywin_load issues the row loads (clamped into the
// slice, so unconditional); ywin_stage works out the rows' segment masks and stores
// the keys transposed into kt[line][WIN_R - H + r]; ywin_find lets the lanes search
// their lines, returns whether some voxel did not resolve and ends with a barrier, so
// kt may be reused.
template <int H>
struct YRows {
  static constexpr int NR = YB + 2 * H;
  static constexpr int RPW = NR / YW_WAVES;
  static_assert(NR % YW_WAVES == 0 && H <= WIN_R, "whole rows per wave");
  uint32_t v[RPW];
};

template <int H>
__device__ __forceinline__ void ywin_load(YRows<H>& rows, const __amdgpu_buffer_rsrc_t zr, uint32_t xl, uint32_t y0,
                                          uint32_t dx, uint32_t dy, uint32_t w)
{
  constexpr int RPW = YRows<H>::RPW;
  const int32_t ybase = (int32_t)y0 - H + (int32_t)w;
  if ((int32_t)y0 >= H && y0 + YB + H <= dy) {  // every staged row inside the map (block-uniform)
    const uint32_t off0 = (uint32_t)ybase * dx * 2, step = YW_WAVES * dx * 2;
#pragma unroll
    for (int i = 0; i < RPW; ++i)
      rows.v[i] = __builtin_amdgcn_raw_buffer_load_b16(zr, (int)(xl * 2), (int)(off0 + (uint32_t)i * step), 0);
  } else {  // rows outside the map read a row inside it (their keys are KEY_NONE)
#pragma unroll
    for (int i = 0; i < RPW; ++i) {
      const int32_t yc = min(max(ybase + YW_WAVES * i, 0), (int32_t)dy - 1);
      rows.v[i] = __builtin_amdgcn_raw_buffer_load_b16(zr, (int)(xl * 2), (int)((uint32_t)yc * dx * 2), 0);
    }
  }
}

template <int H>
__device__ __forceinline__ void ywin_stage(uint32_t* kt, const YRows<H>& rows, const uint32_t* __restrict__ colm,
                                           uint32_t x0, uint32_t y0, uint32_t z, uint32_t dx, uint32_t dy,
                                           uint32_t lane, uint32_t w)
{
  constexpr int NR = YRows<H>::NR, RPW = YRows<H>::RPW;
  const int32_t ybase = (int32_t)y0 - H + (int32_t)w;
  const int32_t zc = (int32_t)(z & 1023u);  // z < 1024: tells the compiler dz fits a 24-bit product
  // segment bits of the staged rows (y0 is a multiple of 32): rm[k] bit b = row 32 k + b
  constexpr int S = (32 - H % 32) % 32;  // (y0 - H) mod 32
  constexpr int NW = (NR + 31) / 32;
  const int32_t base = ((int32_t)y0 - H) >> 5;  // floor((y0 - H) / 32)
  const int32_t nyw = (int32_t)((dy + 31) / 32);
  uint32_t cw[NW + 1], ok[NW];
#pragma unroll
  for (int j = 0; j <= NW; ++j) {
    const int32_t wi = base + j;
    cw[j] = dx % 64 != 0 ? 0xFFFFFFFFu : wi >= 0 && wi < nyw ? colm[(uint32_t)wi * (dx / 64) + x0 / 64] : 0u;
  }
  // staged rows that hold sites: r in [rlo, rhi) (y inside the map) with a stored
  // segment; ok[k] bit b = row 32 k + b + w, i.e. shifted by the wave's row offset so
  // that row w + YW_WAVES i is bit (YW_WAVES i) % 32 of word (YW_WAVES i) / 32
  const int32_t rlo = H - (int32_t)y0, rhi = (int32_t)dy - (int32_t)y0 + H;
#pragma unroll
  for (int k = 0; k < NW; ++k) {
    const uint32_t seg = S == 0 ? cw[k] : (cw[k] >> S) | (cw[k + 1] << ((32 - S) % 32));
    const int32_t a = min(max(rlo - 32 * k, 0), 32), b = min(max(rhi - 32 * k, 0), 32);
    const uint32_t below_b = b >= 32 ? 0xFFFFFFFFu : (1u << b) - 1u, below_a = a >= 32 ? 0xFFFFFFFFu : (1u << a) - 1u;
    ok[k] = (seg & below_b & ~below_a) >> w;
  }
#pragma unroll
  for (int i = 0; i < RPW; ++i) {
    const uint32_t r = w + YW_WAVES * i;
    const int32_t y = ybase + YW_WAVES * i;
    const bool rowok = (ok[(YW_WAVES * i) / 32] >> ((YW_WAVES * i) % 32)) & 1u;  // block-uniform
    const uint32_t vi = rows.v[i];
    // y key (SH = 20): min((z_site - z)^2, 2047) << 20 | y << 10 | z_site
    const int32_t dzv = (int32_t)vi - zc;
    const uint32_t c = min((uint32_t)__mul24(dzv, dzv), 2047u);
    // rows above the map (y < 0, KEY_NONE rows) keep only 10 bits of y, so every key
    // stays below 2^31
    const uint32_t key = (c << 20) | (((uint32_t)y << 10) & 0xFFC00u) | vi;
    // a row without sites (uniform) ORs KEY_NONE (bits 0-30) in, with no branch: every
    // such row reads KEY_NONE, every other key is unchanged.  A column without a site
    // (ZNONE) keeps its key: its cost saturates at 2047, above the cost of any site
    // that resolves a voxel (below (WIN_R + 1)^2), so it never decides one.  Lanes past
    // dx stage lines that no active lane searches.
    const uint32_t none_row = rowok ? 0u : KEY_NONE;
    kt[lane * YS + (WIN_R - H) + r] = key | none_row;
  }
}

template <int H>
__device__ __forceinline__ bool ywin_find(uint32_t* kt, uint32_t x0, uint32_t y0, uint32_t dx, uint32_t dy,
                                          uint32_t lane, uint32_t w, uint32_t* res)
{
  __syncthreads();  // every row staged
  const bool xin = x0 + lane < dx;
  bool fail = false;
#pragma unroll
  for (int i = 0; i < YW_RPS; ++i) {
    const uint32_t j = w * YW_RPS + (uint32_t)i;  // output row y0 + j (uniform over the wave)
    bool okl = true;
    res[i] = win_search<20>(kt + lane * YS + j, xin && y0 + j < dy, okl, H);
    fail = fail || !okl;
  }
  return __syncthreads_or(fail) != 0;  // also: every search is done before kt is reused
}

// the row block's winners (keys holding their sites) as relative sites (yrel_win):
// a wave's lanes hold whole rows, so each row is one coalesced store
__device__ __forceinline__ void ywin_store(const uint32_t* res, uint16_t* __restrict__ outl, uint32_t x0,
                                           uint32_t y0, uint32_t z, uint32_t dx, uint32_t dy, uint32_t lane,
                                           uint32_t w)
{
  const uint32_t x = x0 + lane, yw = y0 + w * YW_RPS;
  const __amdgpu_buffer_rsrc_t orr = slice_rsrc(outl, dx * dy * 2);
  const uint32_t c = yrel_base(yw, z), so = yw * dx * 2;  // row i: c - i, so + i * dx * 2
  if (y0 + YB <= dy && x0 + TB <= dx) {  // block-uniform: every row and lane inside the map
#pragma unroll
    for (int i = 0; i < YW_RPS; ++i)
      __builtin_amdgcn_raw_buffer_store_b16((uint16_t)yrel_win(res[i], c - (uint32_t)i), orr, (int)(x * 2),
                                            (int)(so + (uint32_t)i * dx * 2), 0);
  } else {
#pragma unroll
    for (int i = 0; i < YW_RPS; ++i)
      if (yw + (uint32_t)i < dy && x < dx)
        __builtin_amdgcn_raw_buffer_store_b16((uint16_t)yrel_win(res[i], c - (uint32_t)i), orr, (int)(x * 2),
                                              (int)(so + (uint32_t)i * dx * 2), 0);
  }
}

// one block's row blocks y0 and y0 + YB (when inside the map) with halo H: the second
// block's rows are loaded while the first one is searched.  A row block that does not
// resolve is staged again with halo WIN_R; one that fails then leaves the group to the
// stack kernel.  Returns whether the group was left to it.
template <int H>
__device__ __forceinline__ bool ywin_pair(uint32_t* kt, const __amdgpu_buffer_rsrc_t zr,
                                          const uint32_t* __restrict__ colm, uint16_t* __restrict__ outl,
                                          uint32_t x0, uint32_t y0, uint32_t z, uint32_t dx, uint32_t dy,
                                          uint32_t lane, uint32_t w)
{
  const uint32_t xl = x0 + lane < dx ? x0 + lane : dx - 1;
  const bool two = y0 + YB < dy;  // block-uniform
  uint32_t res[YW_RPS];
  YRows<H> a, b;
  ywin_load<H>(a, zr, xl, y0, dx, dy, w);
  ywin_stage<H>(kt, a, colm, x0, y0, z, dx, dy, lane, w);
  if (two) ywin_load<H>(b, zr, xl, y0 + YB, dx, dy, w);  // in flight during the first search
  const bool fa = ywin_find<H>(kt, x0, y0, dx, dy, lane, w, res);
  if (!fa) ywin_store(res, outl, x0, y0, z, dx, dy, lane, w);
  bool fb = false;
  if (two) {
    ywin_stage<H>(kt, b, colm, x0, y0 + YB, z, dx, dy, lane, w);
    fb = ywin_find<H>(kt, x0, y0 + YB, dx, dy, lane, w, res);
    if (!fb) ywin_store(res, outl, x0, y0 + YB, z, dx, dy, lane, w);
  }
  // a row block that did not resolve is staged again with halo WIN_R, after both
  // (so the rare retry does not add to the registers live above)
  for (int k = 0; k < 2; ++k) {
    if (!(k == 0 ? fa : fb)) continue;
    if (H == WIN_R) return true;
    const uint32_t yk = y0 + (uint32_t)k * YB;
    YRows<WIN_R> full;
    ywin_load<WIN_R>(full, zr, xl, yk, dx, dy, w);
    ywin_stage<WIN_R>(kt, full, colm, x0, yk, z, dx, dy, lane, w);
    if (ywin_find<WIN_R>(kt, x0, yk, dx, dy, lane, w, res)) return true;
    ywin_store(res, outl, x0, yk, z, dx, dy, lane, w);
  }
  return false;
}

__global__ void __launch_bounds__(TB * YW_WAVES, 8) k_pba_y_win(const uint16_t* __restrict__ zs,
                                                             uint16_t* __restrict__ out,
                                                             const uint32_t* __restrict__ tflags,
                                                             const uint32_t* __restrict__ colm,
                                                             unsigned long long* __restrict__ xflags,
                                                             uint32_t* __restrict__ yfs,
                                                             const uint32_t* __restrict__ xcnt, uint32_t dx,
                                                             uint32_t dy, UDiv dzd, UDiv nybd, uint32_t ngroups)
{
  __shared__ uint32_t kt[TB * YS];
  const uint32_t xcd = blockIdx.x % 8, seq = blockIdx.x / 8;
  uint32_t jb, z;
  const uint32_t g = nybd.div(seq, jb) * 8 + xcd;
  if (g >= ngroups) return;  // uniform over the block
  const uint32_t dz = dzd.d, x0 = dzd.div(g, z) * TB, y0 = jb * 2 * YB;  // g = the stack kernel's block index
  const uint32_t lane = threadIdx.x % TB, w = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x / TB));
  const uint32_t ntl = (dy + TYR - 1) / TYR;
  const uint32_t all = ntl >= 32 ? 0xFFFFFFFFu : (1u << ntl) - 1u;
  const uint32_t tf = tflags[x0 / 64] & all;
  if (tf == 0) return;  // no obstacle in these columns: no line keeps a site, nothing is read later
  // halo from the density (see above): the smallest multiple of WIN_CHECK with
  // H^2 rho >= 2.5^2, at most WIN_R (in float: it only picks the halo tried first)
  const uint64_t cols = (uint64_t)(dx - x0 < (uint32_t)TB ? dx - x0 : (uint32_t)TB) * dy * dz;
  const uint32_t nob = xcnt[x0 / 64];
  const float colsf = (float)(dx - x0 < (uint32_t)TB ? dx - x0 : (uint32_t)TB) * (float)dy * (float)dz;
  const float hf = __builtin_amdgcn_sqrtf(6.25f * colsf * __builtin_amdgcn_rcpf((float)(nob ? nob : 1u)));
  int hh = (int)__builtin_ceilf(hf * (1.0f / WIN_CHECK)) * WIN_CHECK;
  hh = __builtin_amdgcn_readfirstlane(hh < WIN_CHECK ? WIN_CHECK : hh > WIN_R ? WIN_R : hh);
  // row tiles without an obstacle (far voxels), or too few obstacles: left to the stack kernel
  if (tf != all || (uint64_t)nob * 256 < cols) {  // the same for every z of the x range
    if (threadIdx.x == 0) yfs[dz + g] = 1u;
    return;
  }
  // (a block of a slice that another block has already flagged still runs: the stack
  // kernel rewrites the whole slice after this kernel, and checking first would cost
  // every block a barrier and a load round)
  const __amdgpu_buffer_rsrc_t zr = slice_rsrc(zs + (uint64_t)z * dx * dy, dx * dy * 2);
  uint16_t* __restrict__ outl = out + (uint64_t)z * dx * dy;
  bool left;
  switch (hh) {  // block-uniform
#define GVX_YWIN_CASE(HV) \
  case HV: left = ywin_pair<HV>(kt, zr, colm, outl, x0, y0, z, dx, dy, lane, w); break;
    GVX_YWIN_CASE(4) GVX_YWIN_CASE(8) GVX_YWIN_CASE(12) GVX_YWIN_CASE(16)
    GVX_YWIN_CASE(20) GVX_YWIN_CASE(24) GVX_YWIN_CASE(28)
#undef GVX_YWIN_CASE
    default: left = ywin_pair<WIN_R>(kt, zr, colm, outl, x0, y0, z, dx, dy, lane, w);
  }
  if (left) {
    if (threadIdx.x == 0) yfs[z] = 1u;
    return;
  }
  // every line of the range kept a site (this row block resolved with a site in every
  // line): its x tiles are read by the x pass
  if (threadIdx.x == 0) {
    unsigned long long f = 0;
#pragma unroll
    for (int k = 0; k < TB / TX; ++k)
      if (x0 + k * TX < dx) f |= 1ull << (x0 / TX + k);
    atomicOr(xflags + z, f);
  }
}

// x pass, phase 2: lines (y, z), lanes = consecutive y.  Stack entry = x |
// (y_site | z_site << 10) << 10, i.e. the packed DistanceVoxel of the site.
// Tiles of 64 lines x TX positions move through LDS so global loads are
// contiguous row segments.
// Tiles of TX positions whose y lines of this slice kept no site (xflags, from the
// y pass) hold only uninitialised voxels: they are neither loaded nor walked.
__global__ void __launch_bounds__(TB) k_pba_x(const uint32_t* __restrict__ map, const uint16_t* __restrict__ yr,
                                              const uint32_t* __restrict__ yfs, uint32_t* __restrict__ stk,
                                              uint32_t* __restrict__ cnt,
                                              const unsigned long long* __restrict__ xflags,
                                              const uint32_t* __restrict__ xfb, uint32_t dx,
                                              uint32_t dy, uint32_t dz, uint32_t uninit)
{
  constexpr int LPR = 64 / TX;  // loader: rows covered by one pass of the wave
  if (xfb && xfb[blockIdx.x] == 0u) return;  // the windowed kernel resolved this group
  constexpr int NLD = TB / LPR;   // loader passes per tile
  __shared__ uint32_t tile[TB][TX + 1];
  __shared__ uint32_t ring[RK][TB];
  __shared__ uint64_t occ[NLD];   // per loader pass: which loaded values hold a site
  // block coordinates as scalars (the buffer descriptor below must be wave-uniform)
  // consecutive blocks take consecutive z with the same y range (every XCD gets
  // every y range); block coordinates as scalars for the buffer descriptor
  const uint32_t z = (uint32_t)__builtin_amdgcn_readfirstlane((int)(blockIdx.x % dz));
  const uint32_t y0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)((blockIdx.x / dz) * TB));
  const uint32_t lane = threadIdx.x, y = y0 + lane;
  const bool active = y < dy;
  const int32_t yi = (int32_t)y, zi = (int32_t)z;
  auto cost = [&](uint32_t yz) -> int32_t { return sq24((int32_t)(yz & 0x3FFu) - yi) + sq24((int32_t)(yz >> 10) - zi); };
  // loader mapping: pass k covers rows LPR k + lane / TX, column lane % TX
  const uint32_t lr = lane / TX, lc = lane % TX;
  const uint32_t ntiles = (dx + TX - 1) / TX;
  LineStack st{ring, stk + ((uint64_t)z * dy + (active ? y : 0)) * dx, lane, 0, 0};
  int32_t tx = 0, tc = 0, sx = 0, sc = 0;
  uint32_t tyz = 0, syz = 0, e3 = 0;  // top, second, third (entry) of the stack
  Bis bst{0, 1};
  uint32_t pre[NLD];
  const uint32_t* __restrict__ slice = map + (uint64_t)z * dy * dx;
  const uint16_t* __restrict__ rslice = yr + (uint64_t)z * dy * dx;
  // the tile's y lines went to the stack kernel (their slice, yfs[z], or their 64-column
  // group, yfs[dz + group]): packed sites in map, else relative sites in yr (uniform
  // over the block)
  const bool pks = yfs[z] != 0u;
  auto packed = [&](uint32_t x0) {  // uniform (a scalar branch keeps the tile loads together)
    return pks || __builtin_amdgcn_readfirstlane((int)yfs[dz + (x0 / 64) * dz + z]) != 0;
  };
  // one lane offset plus the row step; in the last block of a line
  // range (rows < TB) rows past dy read the next slice or, after the last slice, the
  // TB rows of padding behind both buffers (their lanes are inactive)
  auto load_tile = [&](uint32_t x0) {
    const uint32_t xc = x0 + lc < dx ? x0 + lc : dx - 1;
    const uint32_t vo = (y0 + lr) * dx + xc;
    if (!packed(x0)) {
      const __amdgpu_buffer_rsrc_t rr = slice_rsrc(rslice, (dx * dy + TB * dx) * 2);
#pragma unroll
      for (int k = 0; k < NLD; ++k)
        pre[k] = __builtin_amdgcn_raw_buffer_load_b16(rr, (int)((vo + LPR * k * dx) * 2), 0, 0);
    } else {
      const __amdgpu_buffer_rsrc_t mr = slice_rsrc(slice, (dx * dy + TB * dx) * 4);
#pragma unroll
      for (int k = 0; k < NLD; ++k)
        pre[k] = __builtin_amdgcn_raw_buffer_load_b32(mr, (int)((vo + LPR * k * dx) * 4), 0, 0);
    }
  };
  unsigned long long todo = xflags[z] & (ntiles >= 64 ? ~0ull : (1ull << ntiles) - 1ull);
  if (todo) load_tile((uint32_t)__builtin_ctzll(todo) * TX);
  while (todo) {
    const uint32_t t = (uint32_t)__builtin_ctzll(todo);
    todo &= todo - 1ull;
    const uint32_t x0 = t * TX;
    const uint32_t cols = dx - x0 < (uint32_t)TX ? dx - x0 : (uint32_t)TX;
    __syncthreads();
    if (!packed(x0)) {  // relative sites: every one is a site (rows >= dy: inactive lanes)
      const uint32_t xc = x0 + lc < dx ? x0 + lc : dx - 1;
#pragma unroll
      for (int k = 0; k < NLD; ++k) pre[k] = yrel_site(pre[k], xc, y0 + LPR * k + lr, z);
    }
#pragma unroll
    for (int k = 0; k < NLD; ++k) {
      tile[LPR * k + lr][lc] = pre[k];
      const uint64_t b = __ballot(pre[k] != uninit);
      if (lane == 0) occ[k] = b;
    }
    __syncthreads();
    load_tile(todo ? (uint32_t)__builtin_ctzll(todo) * TX : x0);  // the next tile (the last one reloads itself)
    if (active) {
      uint32_t m = (uint32_t)((occ[lane / LPR] >> (TX * (lane % LPR))) & ((1ull << TX) - 1ull));
      if (cols < (uint32_t)TX) m &= (1u << cols) - 1u;
      // the next site's value is read from LDS while the current one is processed
      uint32_t lx = m ? (uint32_t)__builtin_ctz(m) : 0u;
      uint32_t e = tile[lane][lx];
      while (m) {
        m &= m - 1u;
        const uint32_t lxn = m ? (uint32_t)__builtin_ctz(m) : lx;
        const uint32_t en = tile[lane][lxn];
        const int32_t x = (int32_t)(x0 + lx);
        const uint32_t yz = e >> 10;
        const int32_t c = cost(yz);
        if (st.n >= 1) {
          Bis btn;
          btn.set(tx, tc, x, c);
          while (st.n >= 2 && middle_empty(bst, btn)) {
            // pop the top: second and third move up, the new third is fetched
            st.pop();
            tx = sx;
            tyz = syz;
            tc = sc;
            sx = (int32_t)(e3 & 0x3FFu);
            syz = e3 >> 10;
            sc = cost(syz);
            bst.set(sx, sc, tx, tc);
            e3 = st.at3(st.n >= 3 ? st.n - 3 : 0);
            btn.set(tx, tc, x, c);
          }
          bst = btn;
        }
        st.push((uint32_t)x | (yz << 10));
        e3 = (uint32_t)sx | (syz << 10);
        sx = tx;
        syz = tyz;
        sc = tc;
        tx = x;
        tyz = yz;
        tc = c;
        lx = lxn;
        e = en;
      }
    }
  }
  // the final stack goes to the line's global array (bottom first) for the colouring kernel
  if (active) {
    for (uint32_t d = st.lo; d < st.n; ++d) st.g[d] = ring[d % RK][lane];
    cnt[(uint64_t)z * dy + y] = st.n;
  }
}

// floor(N / D) for D > 0 and |N| < 2^22: the float quotient is within 1 of N / D,
// and one integer step corrects its floor
__device__ __forceinline__ int32_t floor_div(int32_t N, int32_t D)
{
  int32_t q = (int32_t)floorf((float)N / (float)D);
  if (__mul24(q, D) > N) --q;
  else if (__mul24(q + 1, D) <= N) ++q;
  return q;
}

// inclusive prefix maximum over the 64 lanes of a wave: DPP row shifts inside
// each row of 16 lanes (lanes shifted in from outside the row read 0), then the
// totals of lanes 15, 31 and 47 carried into the rows above
__device__ __forceinline__ uint32_t wave_prefix_max(uint32_t v, uint32_t lane)
{
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true));  // row_shr:1
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true));  // row_shr:2
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true));  // row_shr:4
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true));  // row_shr:8
  const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)v, 15);
  if (lane >= 16) v = max(v, r0);
  const uint32_t r1 = (uint32_t)__builtin_amdgcn_readlane((int)v, 31);
  if (lane >= 32) v = max(v, r1);
  const uint32_t r2 = (uint32_t)__builtin_amdgcn_readlane((int)v, 47);
  if (lane >= 48) v = max(v, r2);
  return v;
}

// x pass, colouring: one wave per line (y, z), XC_LPW lines per wave one after
// another, XC_LINES waves per block; lane l takes positions 4 l .. 4 l + 3 of each
// 256-position chunk, so a wave stores 1 KiB of contiguous row per instruction.
// Stack entry k >= 1 owns the positions p with p * D_k > N_k, i.e. p >= start_k =
// floor(N_k / D_k) + 1, for the bisector (N_k, D_k) of entries k - 1 and k.  Phase
// 2 leaves these bisectors strictly increasing in k, so the owner of p is the
// largest k with start_k <= p (0 if none) -- the entry at which the sequential
// top-down colouring stops -- i.e. the prefix maximum over positions of
// mark[start_k] = k.  LDS per wave: mark and entries, dx rounded up to 256 each.
// Lines of groups the windowed kernel resolved are skipped (already written).
// Units of XC_LINES * XC_LPW consecutive lines are dealt to a bounded grid (unit
// u to block u % gridDim.x): a block reads the group flags of all its units with
// one load per lane first, so a field the windowed kernel resolved costs the
// dispatch of that grid and one load, not one block per unit.
constexpr int XC_LINES = 4;
constexpr int XC_LPW = 4;
__host__ __device__ constexpr uint32_t xc_pad(uint32_t dx) { return (dx + 255u) & ~255u; }
__device__ __forceinline__ void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__global__ void __launch_bounds__(TB * XC_LINES) k_pba_x_colour(uint32_t* __restrict__ map,
                                                                const uint32_t* __restrict__ stk,
                                                                const uint32_t* __restrict__ cnt,
                                                                const uint32_t* __restrict__ xfb, uint32_t dx,
                                                                UDiv dyd, uint32_t dz, uint32_t uninit)
{
  extern __shared__ uint32_t xc_lds[];
  const uint32_t dy = dyd.d;
  const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x / TB)), lane = threadIdx.x % TB;
  const uint32_t pad = xc_pad(dx);
  uint32_t* mark = xc_lds + (size_t)w * 2 * pad;
  uint32_t* ent = mark + pad;
  const bool vec = (dx & 3u) == 0;  // rows 16-byte aligned: whole uint4 stores
  constexpr uint32_t UL = XC_LINES * XC_LPW;  // lines per unit
  const uint32_t nlines = dy * dz, units = (nlines + UL - 1) / UL;  // dy <= 1024: below 2^32
  // this block's units u = blockIdx.x + i gridDim.x, i < 64 (the grid covers units / 64):
  // lane i tells whether unit i may hold a flagged line (with dy % UL == 0 a unit lies
  // in one group; otherwise every unit is walked and its lines checked one by one)
  unsigned long long todo;
  {
    const uint64_t u = blockIdx.x + (uint64_t)lane * gridDim.x;
    bool f = false;
    if (u < units) {
      uint32_t y0;
      const uint32_t z0 = dyd.div((uint32_t)u * UL, y0);
      f = !xfb || dy % UL != 0 || xfb[(uint64_t)(y0 / TB) * dz + z0] != 0u;
    }
    todo = wave_ballot(f);  // the same in every wave of the block
  }
  while (todo) {  // no block-wide barrier in here
    const uint32_t ubase = (blockIdx.x + (uint32_t)__builtin_ctzll(todo) * gridDim.x) * UL;
    todo &= todo - 1ull;
    for (uint32_t li = 0; li < (uint32_t)XC_LPW; ++li) {
      const uint32_t ln = ubase + w * XC_LPW + li;
      if (ln >= nlines) return;
      uint32_t ly;
      const uint32_t lz = dyd.div(ln, ly);
      if (xfb && xfb[(uint64_t)(ly / TB) * dz + lz] == 0u) continue;
      const uint32_t* __restrict__ s = stk + (uint64_t)ln * dx;
      // the first 64 entries are loaded together with the count (slots past the
      // count hold stale values and are not used)
      const uint32_t n = cnt[ln];
      const uint32_t e0 = s[lane < dx ? lane : 0], a0 = s[lane >= 1 && lane <= dx ? lane - 1 : 0];
      uint32_t* __restrict__ row = map + (uint64_t)ln * dx;
      if (n == 0) {
        for (uint32_t p = lane; p < dx; p += TB) row[p] = uninit;
        continue;
      }
      wave_lds_sync();  // the previous line's reads of mark / ent are done
      for (uint32_t p = 4 * lane; p < pad; p += 4 * TB) *reinterpret_cast<uint4*>(mark + p) = make_uint4(0, 0, 0, 0);
      const int32_t yi = (int32_t)ly, zi = (int32_t)lz;
      auto cost = [&](uint32_t e) -> int32_t {
        return sq24((int32_t)((e >> 10) & 0x3FFu) - yi) + sq24((int32_t)(e >> 20) - zi);
      };
      wave_lds_sync();
      for (uint32_t k = lane; k < n; k += TB) {
        const uint32_t e = k == lane ? e0 : s[k];
        ent[k] = e;
        if (k >= 1) {
          const uint32_t a = k == lane ? a0 : s[k - 1];
          Bis b;
          b.set((int32_t)(a & 0x3FFu), cost(a), (int32_t)(e & 0x3FFu), cost(e));
          const int32_t st = floor_div(b.N, b.D) + 1;
          if (st < (int32_t)dx) atomicMax(&mark[st > 0 ? st : 0], k);
        }
      }
      wave_lds_sync();
      uint32_t carry = 0;
      for (uint32_t c0 = 0; c0 < dx; c0 += 4 * TB) {
        const uint32_t p = c0 + 4 * lane;
        const uint4 mk = *reinterpret_cast<const uint4*>(mark + p);
        uint32_t o0 = mk.x, o1 = max(o0, mk.y), o2 = max(o1, mk.z), o3 = max(o2, mk.w);
        const uint32_t incl = wave_prefix_max(o3, lane);
        uint32_t base = __shfl_up(incl, 1);
        base = max(lane == 0 ? 0u : base, carry);
        carry = max(carry, (uint32_t)__builtin_amdgcn_readlane((int)incl, TB - 1));
        o0 = max(o0, base);
        o1 = max(o1, base);
        o2 = max(o2, base);
        o3 = max(o3, base);
        const uint4 v = make_uint4(ent[o0], ent[o1], ent[o2], ent[o3]);
        if (vec) {
          if (p < dx) *reinterpret_cast<uint4*>(row + p) = v;
        } else {
          if (p < dx) row[p] = v.x;
          if (p + 1 < dx) row[p + 1] = v.y;
          if (p + 2 < dx) row[p + 2] = v.z;
          if (p + 3 < dx) row[p + 3] = v.w;
        }
      }
    }
  }
}

// x pass, windowed: a wave takes XW_RPW rows (y, z) one after another.  The row's
// y-pass entries (relative sites from yr, or packed ones from yb in the slices the y
// stack kernel redid) and their keys (cost (y - y_site)^2 + (z - z_site)^2, position x)
// are staged in LDS, the keys with WIN_R KEY_NONE entries on each side; lane l
// searches positions l + 64 j.  The output at p is the winner's entry, i.e. its
// packed obstacle.  Rows are written only when all of them resolved: a flagged
// group is redone from the untouched y-pass buffer by the proximate-point kernels.
constexpr int XW_WAVES = 4;
constexpr int XW_RPW = 8;  // rows per wave, one after another
// NJ: positions per lane (dx <= 64 NJ)
template <int NJ>
__global__ void __launch_bounds__(TB * XW_WAVES) k_pba_x_win(const uint32_t* __restrict__ yb,
                                                             const uint16_t* __restrict__ yr,
                                                             const uint32_t* __restrict__ yfs,
                                                             uint32_t* __restrict__ map,
                                                             const unsigned long long* __restrict__ xflags,
                                                             uint32_t* __restrict__ xfb, uint32_t dx, UDiv dyd,
                                                             uint32_t dz, uint32_t uninit)
{
  extern __shared__ uint32_t xw_lds[];
  const uint32_t dy = dyd.d;
  const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x / TB)), lane = threadIdx.x % TB;
  const uint32_t kstride = dx + 2 * WIN_R;
  uint32_t* key = xw_lds + (size_t)w * (kstride + dx);  // key[WIN_R + p]
  uint32_t* ent = key + kstride;                        // ent[p]
  const uint32_t ntiles = (dx + TX - 1) / TX;
  const unsigned long long allt = ntiles >= 64 ? ~0ull : (1ull << ntiles) - 1ull;
  const bool full = dx == (uint32_t)(TB * NJ);  // every lane's positions lie in the row
  if (lane < (uint32_t)WIN_R) {
    key[lane] = KEY_NONE;
    key[WIN_R + dx + lane] = KEY_NONE;
  }
  // lines (y, z) number below 2^32 (dy <= 1024); the wave's rows follow one another,
  // so (y, z) is divided out once and then stepped
  const uint32_t nlines = dy * dz, line0 = ((uint32_t)blockIdx.x * XW_WAVES + w) * XW_RPW;
  uint32_t y, yl;
  const uint32_t zf = dyd.div(line0 < nlines ? line0 : 0u, y);
  const uint32_t zl = dyd.div(line0 + XW_RPW - 1 < nlines ? line0 + XW_RPW - 1 : nlines - 1, yl);
  // the wave's rows lie in at most two slices: their x-tile masks are read before any
  // entry, so that no later wait for them also waits for a prefetch
  const unsigned long long tf = line0 < nlines ? xflags[zf] & allt : 0ull, tl = line0 < nlines ? xflags[zl] & allt : 0ull;
  // format of the y output of the wave's (at most two) slices, set by the y stack
  // kernel: 0 = relative sites in yr, 1 = packed sites in yb (the slice, or every
  // 64-column group of it, went to the y stack kernel), 2 = both (its rows are left to
  // the x stack kernels, which read each tile in its own format)
  const uint32_t* __restrict__ yfmt = yfs + dz + (dx + TB - 1) / TB * dz;
  const uint32_t fmtf = line0 < nlines ? yfmt[zf] : 0u, fmtl = line0 < nlines ? yfmt[zl] : 0u;
  // the next row's entries are loaded while the current one is searched (relative
  // sites are expanded once they have arrived); only rows that will be searched
  // (every x tile of their slice kept a site) are read: the others' loads get a
  // zero-size buffer and read nothing.  Lanes past dx read past the row (at most
  // into the padding behind the buffers) and are not used.
  auto fetch = [&](uint32_t ln, uint32_t yln, uint32_t zln, uint32_t* dst) {
    const uint32_t fmt = zln == zf ? fmtf : fmtl;
    const bool want = ln < nlines && ((zln == zf ? tf : tl) == allt) && fmt != 2u;  // uniform over the wave
    const uint64_t so = (uint64_t)zln * dx * dy;
    if (fmt == 1u) {
      const __amdgpu_buffer_rsrc_t r = slice_rsrc(yb + so, want ? (dx * dy + TB * dx) * 4 : 0u);
#pragma unroll
      for (int j = 0; j < NJ; ++j)
        dst[j] = __builtin_amdgcn_raw_buffer_load_b32(r, (int)(lane * 4), (int)((yln * dx + TB * j) * 4), 0);
    } else {
      const __amdgpu_buffer_rsrc_t r = slice_rsrc(yr + so, want ? (dx * dy + TB * dx) * 2 : 0u);
#pragma unroll
      for (int j = 0; j < NJ; ++j)
        dst[j] = __builtin_amdgcn_raw_buffer_load_b16(r, (int)(lane * 2), (int)((yln * dx + TB * j) * 2), 0);
    }
  };
  uint32_t e[NJ], en[NJ];
  uint32_t z = zf;
  fetch(line0, y, z, en);
  for (uint32_t rr = 0; rr < (uint32_t)XW_RPW; ++rr) {  // no block-wide barrier in here
    const uint32_t line = line0 + rr;
    if (line >= nlines) return;
    if (rr > 0) {  // the next row: (y + 1, z) or (0, z + 1)
      const bool wrap = y + 1 == dy;
      y = wrap ? 0u : y + 1;
      z = wrap ? z + 1 : z;
    }
    const uint32_t fmt = z == zf ? fmtf : fmtl;
    const bool pk = fmt == 1u;
#pragma unroll
    for (int j = 0; j < NJ; ++j)  // (waits for the prefetch) then the next row's
      e[j] = pk ? en[j] : yrel_site(en[j], lane + TB * j, y, z);
    fetch(rr + 1 < (uint32_t)XW_RPW ? line + 1 : nlines, y + 1 == dy ? 0u : y + 1, y + 1 == dy ? z + 1 : z, en);
    const unsigned long long tiles = z == zf ? tf : tl;
    const uint64_t grp = (uint64_t)(y / TB) * dz + z;
    uint32_t* __restrict__ row = map + (uint64_t)line * dx;
    bool search = false;
    if (tiles == 0) {  // no line of the slice kept a site
      for (uint32_t p = lane; p < dx; p += TB) row[p] = uninit;
    } else if (tiles != allt || fmt == 2u) {  // x ranges without a site (far voxels), or mixed formats
      if (lane == 0) xfb[grp] = 1u;
    } else {
      search = true;
    }
    if (search) {
      const int32_t yi = (int32_t)y, zi = (int32_t)z;
      // the previous row's reads of key / ent are complete before they are overwritten
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      bool any = false;
      if (!pk) {  // from relative sites: every position of a searched row holds one
        auto stage = [&](int j) {
          const uint32_t p = lane + TB * j;
          const int32_t ey = (int32_t)((e[j] >> 10) & 0x3FFu) - yi, ez = (int32_t)(e[j] >> 20) - zi;
          key[WIN_R + p] = site_key((uint32_t)(sq24(ey) + sq24(ez)), p);
          ent[p] = e[j];
        };
        if (full) {
#pragma unroll
          for (int j = 0; j < NJ; ++j) stage(j);
        } else {
#pragma unroll
          for (int j = 0; j < NJ; ++j)
            if (lane + TB * j < dx) stage(j);
        }
        any = true;
      } else {
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
          const uint32_t p = lane + TB * j;
          if (p < dx) {
            const bool site = e[j] != uninit;
            const int32_t ey = (int32_t)((e[j] >> 10) & 0x3FFu) - yi, ez = (int32_t)(e[j] >> 20) - zi;
            key[WIN_R + p] = site ? site_key((uint32_t)(sq24(ey) + sq24(ez)), p) : KEY_NONE;
            ent[p] = e[j];
            any = any || site;
          }
        }
      }
      if (!wave_ballot(any)) {  // a row without a site stays uninitialised
        for (uint32_t p = lane; p < dx; p += TB) row[p] = uninit;
      } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        bool fail = false;
        uint32_t res[NJ];
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
          res[j] = 0;
          if ((uint32_t)(TB * j) < dx) {  // uniform over the wave
            const uint32_t p = lane + TB * j;
            bool ok = true;
            const uint32_t b = win_search<11>(key + (p < dx ? p : 0), p < dx, ok);
            fail = fail || !ok;
            res[j] = (b >> 1) & 1023u;
          }
        }
        if (wave_ballot(fail)) {
          if (lane == 0) xfb[grp] = 1u;
        } else {
          if (full) {
#pragma unroll
            for (int j = 0; j < NJ; ++j) row[lane + TB * j] = ent[res[j]];
          } else {
#pragma unroll
            for (int j = 0; j < NJ; ++j)
              if (lane + TB * j < dx) row[lane + TB * j] = ent[res[j]];
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------- read-outs
// DistanceVoxel::squaredObstacleDistance (DistanceVoxel.hpp:38-55)
__device__ __forceinline__ int32_t dist_sq(uint32_t raw, uint64_t id, uint32_t dx, uint32_t dy, uint32_t uninit)
{
  if (raw == uninit) return INT32_MAX;
  const int32_t x = (int32_t)(id % dx), y = (int32_t)((id / dx) % dy), z = (int32_t)(id / ((uint64_t)dx * dy));
  const int32_t ex = x - (int32_t)(raw & 0x3FF), ey = y - (int32_t)((raw >> 10) & 0x3FF), ez = z - (int32_t)(raw >> 20);
  return ex * ex + ey * ey + ez * ez;
}

__global__ void k_dist_sq(const uint32_t* __restrict__ map, const uint32_t* __restrict__ idx, uint64_t n, uint32_t dx,
                          uint32_t dy, uint32_t uninit, int32_t* __restrict__ out)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint64_t id = idx ? idx[i] : i;
    out[i] = dist_sq(map[id], id, dx, dy, uninit);
  }
}

__global__ void k_dist_metric(const uint32_t* __restrict__ map, uint64_t n, uint32_t dx, uint32_t dy, uint32_t uninit,
                              float side, float* __restrict__ out)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int32_t d = dist_sq(map[i], i, dx, dy, uninit);
    out[i] = d == INT32_MAX ? 3.402823466e+38f : sqrtf((float)d) * side;
  }
}

// mergeOccupied (DistanceVoxelMap.hpp:109-136 / DistanceVoxelMap.h:92-134)
__global__ void k_merge_occupied(uint32_t* __restrict__ dist, const int8_t* __restrict__ prob, uint64_t n, uint32_t dx,
                                 uint32_t dy, uint32_t ox, uint32_t oy, uint32_t oz, int thr)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if ((int)prob[i] > thr) {
      const uint32_t x = (uint32_t)(i % dx), y = (uint32_t)((i / dx) % dy), z = (uint32_t)(i / ((uint64_t)dx * dy));
      dist[i] = (x + ox) | ((y + oy) << 10) | ((z + oz) << 20);
    }
  }
}

// kernelExactDistances3D (VoxelMapOperations.hpp:896-990): obstacle list staged
// through LDS in tiles; strict < keeps the first minimum in list order.
constexpr int EX_BLOCK = 256;
__global__ void __launch_bounds__(EX_BLOCK) k_exact(uint32_t* __restrict__ map, uint64_t n, uint32_t dx, uint32_t dy,
                                                    uint32_t uninit, const float* __restrict__ obst, uint64_t no,
                                                    float side)
{
  __shared__ int32_t ob[EX_BLOCK][3];
  const uint64_t id = (uint64_t)blockIdx.x * EX_BLOCK + threadIdx.x;
  const bool active = id < n;
  int32_t x = 0, y = 0, z = 0, cur = 0, best = 0;
  int32_t bo[3] = {0, 0, 0};
  bool found = false;
  if (active) {
    x = (int32_t)(id % dx); y = (int32_t)((id / dx) % dy); z = (int32_t)(id / ((uint64_t)dx * dy));
    cur = dist_sq(map[id], id, dx, dy, uninit);
    best = cur;
  }
  for (uint64_t t = 0; t < no; t += EX_BLOCK) {
    __syncthreads();
    const uint64_t k = t + threadIdx.x;
    if (k < no) {
      ob[threadIdx.x][0] = quantize_signed(obst[3 * k], side);
      ob[threadIdx.x][1] = quantize_signed(obst[3 * k + 1], side);
      ob[threadIdx.x][2] = quantize_signed(obst[3 * k + 2], side);
    }
    __syncthreads();
    if (!active || cur == 0) continue;
    const int cnt = (int)((no - t) < (uint64_t)EX_BLOCK ? (no - t) : (uint64_t)EX_BLOCK);
    for (int s = 0; s < cnt; ++s) {
      const int32_t a = ob[s][0], b = ob[s][1], c = ob[s][2];
      if (a == x && b == y && c == z) continue;
      if (a == 1023 || b == 1023 || c == 1023) continue;
      const int32_t ex = x - a, ey = y - b, ez = z - c;
      const int32_t dd = ex * ex + ey * ey + ez * ez;
      if (dd < best) { best = dd; bo[0] = a; bo[1] = b; bo[2] = c; found = true; }
    }
  }
  if (active && found && best < cur) map[id] = (uint32_t)bo[0] | ((uint32_t)bo[1] << 10) | ((uint32_t)bo[2] << 20);
}

struct PbaScratch {
  void* zs = nullptr;     // z pass output, uint16 per voxel
  void* ybuf = nullptr;   // y pass output of the stack kernel, the packed DistanceVoxel per voxel
  void* yrel = nullptr;   // y pass output of the windowed kernel, the relative site per voxel (uint16)
  void* node = nullptr;   // stack spills of the y pass, final stacks of the x pass
  void* cnt = nullptr;    // x pass: stack size per line
  void* flags = nullptr;  // z-pass block flags, x-tile masks per slice, stored row segments, group flags
  uint64_t n = 0, lines = 0, flag_words = 0;
};
static PbaScratch& pba_scratch()
{
  static PbaScratch s;
  return s;
}

// GVX_SYNC_CHECK=1 in the environment: every EDT launch is followed by a stream
// synchronisation, so an execution error is reported with the name of its kernel
static bool pba_sync_check()
{
  static const bool on = [] {
    const char* e = getenv("GVX_SYNC_CHECK");
    return e && e[0] == '1';
  }();
  return on;
}

static int pba_launched(const char* kernel)
{
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && pba_sync_check()) e = hipStreamSynchronize(rt().stream);
  if (e != hipSuccess) return fail(GVX_ERR_HIP, std::string("PBA kernel ") + kernel + ": " + hipGetErrorString(e));
  return GVX_OK;
}

// Elements of padding behind the y-pass outputs: the x stack kernel's tile loads of
// the last line block read up to TB rows of at most 1024 positions past a slice.
constexpr uint64_t PBA_ROW_PAD = (uint64_t)TB * 1024;

// Device buffers of one EDT run (sizes: pba_async, pba_carve).
struct PbaBufs {
  void* zs;     // z pass output, uint16 per voxel
  void* ybuf;   // y pass output of the stack kernel, the packed DistanceVoxel per voxel
  void* yrel;   // y pass output of the windowed kernel, the relative site per voxel (uint16)
  void* node;   // stack spills of the y pass, final stacks of the x pass
  void* cnt;    // x pass: stack size per line
  void* flags;  // z-pass block flags, x-tile masks per slice, stored row segments, group flags
};

// 32-bit words of the flag area for dims (dx, dy, dz)
static uint64_t pba_flag_words(uint32_t dx, uint32_t dy, uint32_t dz)
{
  const uint64_t cwords = (uint64_t)((dy + 31) / 32) * ((dx + 63) / 64);
  const uint64_t xgroups = (uint64_t)((dy + TB - 1) / TB) * dz;
  const uint64_t ygroups = (uint64_t)((dx + TB - 1) / TB) * dz;
  return 64 + 2 * (uint64_t)dz + cwords + dz + ygroups + dz + xgroups + 64;
}

static int pba_launch(MapS* m, const PbaBufs& sc, bool timed);

int pba_async(MapS* m)
{
  const uint32_t dx = m->dims[0], dy = m->dims[1], dz = m->dims[2];
  PbaScratch& sc = pba_scratch();
  {
    const hipError_t e = hipGetLastError();  // an error left by an earlier call is not this call's
    if (e != hipSuccess) return fail(GVX_ERR_HIP, std::string("PBA: pending error before the call: ") + hipGetErrorString(e));
  }
  if (dx > 64u * TX || dy > 32u * TYR) return fail(GVX_ERR_UNSUPPORTED, "PBA: x or y dimension above 1024");
  if (sc.n < m->n) {
    GVX_CHECK(sync());
    dfree(sc.zs);
    dfree(sc.ybuf);
    dfree(sc.yrel);
    dfree(sc.node);
    sc.zs = sc.ybuf = sc.yrel = sc.node = nullptr;
    sc.n = 0;
    GVX_CHECK(dalloc(&sc.zs, m->n * 2));
    GVX_CHECK(dalloc(&sc.ybuf, (m->n + PBA_ROW_PAD) * 4));
    GVX_CHECK(dalloc(&sc.yrel, (m->n + PBA_ROW_PAD) * 2));
    GVX_CHECK(dalloc(&sc.node, m->n * 4));
    sc.n = m->n;
  }
  const uint64_t lines = (uint64_t)dy * dz;
  if (sc.lines < lines) {
    GVX_CHECK(sync());
    dfree(sc.cnt);
    sc.cnt = nullptr;
    sc.lines = 0;
    GVX_CHECK(dalloc(&sc.cnt, lines * 4));
    sc.lines = lines;
  }
  const uint64_t fwords = pba_flag_words(dx, dy, dz);
  if (sc.flag_words < fwords) {
    GVX_CHECK(sync());
    dfree(sc.flags);
    sc.flags = nullptr;
    sc.flag_words = 0;
    GVX_CHECK(dalloc(&sc.flags, fwords * 4));
    sc.flag_words = fwords;
  }
  {
    const hipError_t e = hipGetLastError();  // from freeing or allocating the scratch
    if (e != hipSuccess) return fail(GVX_ERR_HIP, std::string("PBA scratch: ") + hipGetErrorString(e));
  }
  return pba_launch(m, PbaBufs{sc.zs, sc.ybuf, sc.yrel, sc.node, sc.cnt, sc.flags}, true);
}

// The three passes on the library stream over the buffers sc; no allocation.
// timed: named HIP-event scopes around the passes (when profiling is on).
static int pba_launch(MapS* m, const PbaBufs& sc, bool timed)
{
  const uint32_t dx = m->dims[0], dy = m->dims[1], dz = m->dims[2];
  // the skip masks hold 32 row tiles per x range and 64 x tiles per slice
  if (dx > 64u * TX || dy > 32u * TYR) return fail(GVX_ERR_UNSUPPORTED, "PBA: x or y dimension above 1024");
  const uint64_t lines = (uint64_t)dy * dz;
  const uint64_t cwords = (uint64_t)((dy + 31) / 32) * ((dx + 63) / 64);  // stored row segments
  const uint64_t ygroups = (uint64_t)((dx + TB - 1) / TB) * dz;          // y pass: 64 lines (x range, z)
  const uint64_t xgroups = (uint64_t)((dy + TB - 1) / TB) * dz;          // x pass: 64 lines (y range, z)
  const uint64_t fwords = pba_flag_words(dx, dy, dz);
  hipStream_t s = rt().stream;
  uint32_t* map = (uint32_t*)m->data;
  uint16_t* zs = (uint16_t*)sc.zs;
  uint32_t* yb = (uint32_t*)sc.ybuf;
  uint16_t* yr = (uint16_t*)sc.yrel;
  uint32_t* st = (uint32_t*)sc.node;
  uint32_t* tflags = (uint32_t*)sc.flags;
  unsigned long long* xflags = (unsigned long long*)((uint32_t*)sc.flags + 64);
  uint32_t* colm = (uint32_t*)sc.flags + 64 + 2 * (uint64_t)dz;
  // y lines left to the stack kernel: yfs[z] per slice (a window search that did not
  // resolve), then yfs[dz + group] per 64-column group (ranges too sparse for the window),
  // then per slice the format of its y output (set by the stack kernel for the x pass)
  uint32_t* yfs = colm + cwords;
  uint32_t* xfb = yfs + dz + ygroups + dz;  // x groups left to the stack kernels
  uint32_t* xcnt = xfb + xgroups;  // obstacles per 64-column x range
  const uint64_t cols = (uint64_t)dx * dy;
  {
    ProfScope ps(timed ? "pba_z" : nullptr);
    GVX_CHECK(fill_u32(sc.flags, 0u, fwords));
    GVX_CHECK(pba_launched("fill of the flags"));
    dim3 g((unsigned)((cols + 63) / 64));
    const uint32_t u = m->uninit;
#define ZL(W)                                                                                               \
  if (dz == 32 * W)                                                                                         \
    hipLaunchKernelGGL((k_pba_z<W, true>), g, dim3(64 * (W / ZWP)), 0, s, map, zs, tflags, colm, xcnt,       \
                       (uint32_t)TYR, dx, dy, dz, u);                                                       \
  else hipLaunchKernelGGL((k_pba_z<W, false>), g, dim3(64 * (W / ZWP)), 0, s, map, zs, tflags, colm, xcnt,  \
                          (uint32_t)TYR, dx, dy, dz, u)
    if (dz <= 128) { ZL(4); }
    else if (dz <= 256) { ZL(8); }
    else if (dz <= 512) { ZL(16); }
    else if (dz <= 1024) { ZL(32); }
    else {
      hipLaunchKernelGGL(k_pba_z_long, dim3((unsigned)((cols + 255) / 256)), dim3(256), 0, s, map, zs, dx, dy, dz, u);
      GVX_CHECK(fill_u32(tflags, 0xFFFFFFFFu, 64));  // no block flags: every tile is walked
      GVX_CHECK(fill_u32(colm, 0xFFFFFFFFu, cwords));  // and every segment was stored
    }
#undef ZL
    GVX_CHECK(pba_launched("k_pba_z"));
  }
  {
    ProfScope ps(timed ? "pba_y" : nullptr);
    const uint32_t nyb = (dy + 2 * YB - 1) / (2 * YB);  // blocks per group: two row blocks each
    const uint64_t wblocks = (ygroups + 7) / 8 * 8 * nyb;  // blocks of 8 groups per XCD slot
    hipLaunchKernelGGL(k_pba_y_win, dim3((unsigned)wblocks), dim3(TB * YW_WAVES), 0, s, zs, yr, tflags, colm, xflags,
                       yfs, xcnt, dx, dy, udiv(dz), udiv(nyb), (uint32_t)ygroups);
    GVX_CHECK(pba_launched("k_pba_y_win"));
    hipLaunchKernelGGL(k_pba_y, dim3((unsigned)ygroups), dim3(TB), 0, s, zs, yb, st, tflags, colm, xflags, yfs, dx, dy,
                       dz, m->uninit);
    GVX_CHECK(pba_launched("k_pba_y"));
  }
  {
    ProfScope ps(timed ? "pba_x" : nullptr);
    uint32_t* cnt = (uint32_t*)sc.cnt;
    const uint64_t xrows = (uint64_t)XW_WAVES * XW_RPW;
    const dim3 xg((unsigned)((lines + xrows - 1) / xrows));
    const size_t xl = (size_t)(2 * dx + 2 * WIN_R) * 4 * XW_WAVES;
    if (dx <= 256)
      hipLaunchKernelGGL(k_pba_x_win<4>, xg, dim3(TB * XW_WAVES), xl, s, yb, yr, yfs, map, xflags, xfb, dx, udiv(dy), dz, m->uninit);
    else if (dx <= 512)
      hipLaunchKernelGGL(k_pba_x_win<8>, xg, dim3(TB * XW_WAVES), xl, s, yb, yr, yfs, map, xflags, xfb, dx, udiv(dy), dz, m->uninit);
    else
      hipLaunchKernelGGL(k_pba_x_win<16>, xg, dim3(TB * XW_WAVES), xl, s, yb, yr, yfs, map, xflags, xfb, dx, udiv(dy), dz, m->uninit);
    GVX_CHECK(pba_launched("k_pba_x_win"));
    hipLaunchKernelGGL(k_pba_x, dim3((unsigned)xgroups), dim3(TB), 0, s, yb, yr, yfs, st, cnt, xflags, xfb, dx, dy, dz,
                       m->uninit);
    GVX_CHECK(pba_launched("k_pba_x"));
    const uint64_t clines = (uint64_t)XC_LINES * XC_LPW, units = (lines + clines - 1) / clines;
    const uint64_t cgrid = std::max<uint64_t>((units + 63) / 64, std::min<uint64_t>(units, 4096));
    hipLaunchKernelGGL(k_pba_x_colour, dim3((unsigned)cgrid), dim3(TB * XC_LINES),
                       (size_t)xc_pad(dx) * 8 * XC_LINES, s, map, st, cnt, xfb, dx, udiv(dy), dz, m->uninit);
    GVX_CHECK(pba_launched("k_pba_x_colour"));
  }
  return GVX_OK;
}

// Workspace of gvx_dist_pba_noalloc: a fixed device array, so that call allocates
// nothing.  In host code `scratch` is only a symbol; its device address (one per
// device, taken once at gvx_init with the device current) is what kernels get.
constexpr size_t PBA_NOALLOC_BYTES = (size_t)32 << 20;
static __device__ __align__(256) uint8_t scratch[PBA_NOALLOC_BYTES];
static struct {
  int device = -1;         // device the address below belongs to
  void* ptr = nullptr;     // device address of scratch there
  std::string error;       // why it could not be taken, if it could not
} g_noalloc;

int pba_noalloc_setup(int device)
{
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || ptr == nullptr) {
    (void)hipGetLastError();
    g_noalloc.device = -1;
    g_noalloc.ptr = nullptr;
    g_noalloc.error = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return fail(GVX_ERR_HIP, g_noalloc.error);
  }
  g_noalloc.device = device;
  g_noalloc.ptr = ptr;
  g_noalloc.error.clear();
  return GVX_OK;
}

// Carves the run's buffers out of [base, base + cap); false when they do not fit.
static bool pba_carve(const MapS* m, uint8_t* base, size_t cap, PbaBufs* b, size_t* need)
{
  const uint32_t dx = m->dims[0], dy = m->dims[1], dz = m->dims[2];
  auto up = [](uint64_t v) { return (v + 255) & ~(uint64_t)255; };
  const uint64_t o_ybuf = up(m->n * 2), o_yrel = o_ybuf + up((m->n + PBA_ROW_PAD) * 4);
  const uint64_t o_node = o_yrel + up((m->n + PBA_ROW_PAD) * 2);
  const uint64_t o_cnt = o_node + up(m->n * 4);
  const uint64_t o_flags = o_cnt + up((uint64_t)dy * dz * 4);
  const uint64_t total = o_flags + up(pba_flag_words(dx, dy, dz) * 4);
  *need = (size_t)total;
  if (total > cap) return false;
  *b = PbaBufs{base, base + o_ybuf, base + o_yrel, base + o_node, base + o_cnt, base + o_flags};
  return true;
}

}  // namespace gvx

using namespace gvx;

static int check_dist(gvx_map m)
{
  if (!m || !m->data) return fail(GVX_ERR_BAD_ARG, "null map handle");
  if (m->type != GVX_MT_DISTANCE_VOXELMAP) return fail(GVX_ERR_UNSUPPORTED, "not a DistanceVoxelMap");
  return GVX_OK;
}

extern "C" {

int gvx_dist_pba(gvx_map m)
{
  GVX_CHECK(check_dist(m));
  std::lock_guard<std::recursive_mutex> g(m->mu);
  std::lock_guard<std::recursive_mutex> g2(rt().mu);
  {
    ProfScope ps("pba");
    GVX_CHECK(pba_async(m));
  }
  return sync();
}

int gvx_dist_pba_noalloc(gvx_map m)
{
  GVX_CHECK(check_dist(m));
  std::lock_guard<std::recursive_mutex> g(m->mu);
  std::lock_guard<std::recursive_mutex> g2(rt().mu);
  if (g_noalloc.ptr == nullptr || g_noalloc.device != rt().device)
    return fail(GVX_ERR_HIP, "gvx_dist_pba_noalloc: no workspace on this device (" +
                                 (g_noalloc.error.empty() ? std::string("gvx_init did not set it up") : g_noalloc.error) +
                                 ")");
  PbaBufs b;
  size_t need = 0;
  if (!pba_carve(m, (uint8_t*)g_noalloc.ptr, PBA_NOALLOC_BYTES, &b, &need))
    return fail(GVX_ERR_BAD_ARG, "gvx_dist_pba_noalloc: a " + std::to_string(m->dims[0]) + "x" +
                                     std::to_string(m->dims[1]) + "x" + std::to_string(m->dims[2]) + " map needs " +
                                     std::to_string(need) + " bytes of workspace, above the fixed " +
                                     std::to_string(PBA_NOALLOC_BYTES) + "; use gvx_dist_pba");
  {
    const hipError_t e = hipGetLastError();  // an error left by an earlier call is not this call's
    if (e != hipSuccess)
      return fail(GVX_ERR_HIP, std::string("PBA: pending error before the call: ") + hipGetErrorString(e));
  }
  GVX_CHECK(pba_launch(m, b, false));
  return sync();
}

int gvx_dist_merge_occupied(gvx_map dist, gvx_map prob, const uint32_t offset[3], float threshold)
{
  GVX_CHECK(check_dist(dist));
  if (!prob || prob->type != GVX_MT_PROBAB_VOXELMAP) return fail(GVX_ERR_BAD_ARG, "mergeOccupied needs a ProbVoxelMap");
  if (memcmp(dist->dims, prob->dims, sizeof(dist->dims)) != 0)
    return fail(GVX_ERR_DIM_MISMATCH, "mergeOccupied: map dimensions differ");
  std::unique_lock<std::recursive_mutex> g1(dist->mu, std::defer_lock), g2(prob->mu, std::defer_lock);
  std::lock(g1, g2);
  hipLaunchKernelGGL(k_merge_occupied, dim3(grid_for(dist->n, 256, 8192)), dim3(256), 0, rt().stream,
                     (uint32_t*)dist->data, (const int8_t*)prob->data, dist->n, dist->dims[0], dist->dims[1],
                     offset ? offset[0] : 0u, offset ? offset[1] : 0u, offset ? offset[2] : 0u,
                     (int)float_to_probability(threshold));
  GVX_HIP(hipGetLastError());
  return sync();
}

int gvx_dist_exact(gvx_map m, const float* obstacles, uint64_t n, int mem)
{
  GVX_CHECK(check_dist(m));
  if (n && !obstacles) return fail(GVX_ERR_BAD_ARG, "null obstacles");
  std::lock_guard<std::recursive_mutex> g(m->mu);
  std::lock_guard<std::recursive_mutex> g2(rt().mu);
  const void* d = nullptr;
  GVX_CHECK(stage_input(obstacles, n * 12, mem, &d));
  hipLaunchKernelGGL(k_exact, dim3((unsigned)((m->n + EX_BLOCK - 1) / EX_BLOCK)), dim3(EX_BLOCK), 0, rt().stream,
                     (uint32_t*)m->data, m->n, m->dims[0], m->dims[1], m->uninit, (const float*)d, n, m->side);
  GVX_HIP(hipGetLastError());
  return sync();
}

int gvx_dist_squared(gvx_map m, const uint32_t* idx, uint64_t n, int32_t* out, int mem)
{
  GVX_CHECK(check_dist(m));
  if (!out) return fail(GVX_ERR_BAD_ARG, "null out");
  if (!idx) n = m->n;
  if (n == 0) return GVX_OK;
  std::lock_guard<std::recursive_mutex> g(m->mu);
  std::lock_guard<std::recursive_mutex> g2(rt().mu);
  // indices staged in front, results behind
  size_t ib = idx ? ((n * 4 + 255) & ~(size_t)255) : 0;
  void* st = nullptr;
  const uint32_t* didx = nullptr;
  int32_t* dout = out;
  if (idx && mem == GVX_MEM_DEVICE) didx = idx;
  size_t need = (idx && mem == GVX_MEM_HOST ? ib : 0) + (mem == GVX_MEM_HOST ? n * 4 : 0);
  if (need) GVX_CHECK(stage_reserve(need, &st));
  if (idx && mem == GVX_MEM_HOST) {
    GVX_HIP(hipMemcpyAsync(st, idx, n * 4, hipMemcpyHostToDevice, rt().stream));
    didx = (const uint32_t*)st;
  }
  if (mem == GVX_MEM_HOST) dout = (int32_t*)((char*)st + (idx ? ib : 0));
  if (idx) {
    // bounds-check indices on the way: invalid ids are rejected before launch
    std::vector<uint32_t> host;
    const uint32_t* hp = idx;
    if (mem == GVX_MEM_DEVICE) {
      host.resize(n);
      GVX_HIP(hipMemcpyAsync(host.data(), idx, n * 4, hipMemcpyDeviceToHost, rt().stream));
      GVX_CHECK(sync());
      hp = host.data();
    }
    for (uint64_t i = 0; i < n; ++i)
      if (hp[i] >= m->n) return fail(GVX_ERR_BAD_ARG, "voxel index out of range");
  }
  hipLaunchKernelGGL(k_dist_sq, dim3(grid_for(n, 256, 8192)), dim3(256), 0, rt().stream, (const uint32_t*)m->data,
                     didx, n, m->dims[0], m->dims[1], m->uninit, dout);
  GVX_HIP(hipGetLastError());
  if (mem == GVX_MEM_HOST) GVX_HIP(hipMemcpyAsync(out, dout, n * 4, hipMemcpyDeviceToHost, rt().stream));
  return sync();
}

int gvx_dist_metric(gvx_map m, float* out, int mem)
{
  GVX_CHECK(check_dist(m));
  if (!out) return fail(GVX_ERR_BAD_ARG, "null out");
  std::lock_guard<std::recursive_mutex> g(m->mu);
  std::lock_guard<std::recursive_mutex> g2(rt().mu);
  float* dout = out;
  if (mem == GVX_MEM_HOST) {
    void* st = nullptr;
    GVX_CHECK(stage_reserve(m->n * 4, &st));
    dout = (float*)st;
  }
  hipLaunchKernelGGL(k_dist_metric, dim3(grid_for(m->n, 256, 8192)), dim3(256), 0, rt().stream,
                     (const uint32_t*)m->data, m->n, m->dims[0], m->dims[1], m->uninit, m->side, dout);
  GVX_HIP(hipGetLastError());
  if (mem == GVX_MEM_HOST) GVX_HIP(hipMemcpyAsync(out, dout, m->n * 4, hipMemcpyDeviceToHost, rt().stream));
  return sync();
}

}  // extern "C"
