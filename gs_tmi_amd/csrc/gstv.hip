// gstv.hip - MI355X (gfx950, CDNA4) tile voxelizer: binning, forward, adjoint.
//
// Hot path of 3DGSMR's tile voxelizer (gaussian/tile_voxelizer.py:87-162,
// gaussian/cuda_ext/tile_voxelize_cuda_kernel.cu:48-349), written for 64-lane
// wavefronts.  Everything sits behind the C ABI declared in include/gstv.h.
//
// Layout of the work
//   * Binning: per-Gaussian tile counts -> device-wide exclusive scan -> stable
//     LSD radix sort on the tile id (balanced digits of at most 8 bits,
//     wave-level multisplit ranking).  The first pass expands the (tile id,
//     Gaussian id) pairs itself, in Gaussian-major order, from the Gaussians'
//     tile boxes; the last pass writes only the Gaussian ids and merges the
//     per-tile [start, end) ranges.  Stability makes the order inside a tile
//     ascending in Gaussian id.
//   * A per-Gaussian record (64 B: centre, density, quadratic form in voxel
//     units pre-scaled by -log2(e)/2, and the exact axis-aligned extent of the
//     q <= 9 ellipsoid) is built once per call, so the tile kernels gather one
//     cache line per (tile, Gaussian) pair.
//   * Column bound: for a 4x4 (H, W) voxel column the D rows that can hold a
//     voxel of the q <= 9 ellipsoid lie inside the intersection of the D
//     extents of its two projections on (D, H) and (D, W), each cut by the
//     column's strip -- a closed form per strip (clamp, square root).  It
//     culls sub-bricks in the forward and rows in the backward; about half of
//     the box-level work has no voxel inside the ellipsoid.
//   * Forward: one 512-thread block per 8^3 tile (one block per 8 sub-bricks
//     for larger tiles); each wave owns a 4x4x4 sub-brick.  A batch of up to
//     1024 pairs is staged in LDS with the mask of sub-bricks the column bound
//     leaves; every wave compacts its Gaussians (ballot) and evaluates them 16
//     at a time: lane = (Gaussian slot 0..15, W offset 0..3), each lane walks
//     the 4x4 (D, H) positions of its W column, so the quadratic form costs
//     two FMAs per evaluation; 16 complex partial sums per lane are reduced
//     across the 16 Gaussian slots once per tile and every voxel is written
//     once.
//   * Backward: one block per tile; each staged Gaussian walks its live
//     columns (4x4 (H, W) blocks of its box, each with the D rows the column
//     bound leaves) with lane = (Gaussian slot, W offset), in groups of 16
//     Gaussians with equal column counts, keeping its 11 gradient moments in
//     registers; a lane-swap reduce-scatter over the 4 W lanes and a lane
//     permute leave each atomic instruction 4 whole 64-byte moment rows, added
//     with fp32 atomics.
//     A final per-Gaussian kernel maps moments to the reference's outputs
//     (centre, cov_inv, density gradients).

#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <type_traits>

#include "gstv.h"

namespace {

// ------------------------------------------------------------------ common --

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;                      // items per thread in scans
constexpr int kScanTile = kScanThreads * kScanItems;
constexpr int kScanFoldBlocks = 1024;                // scans of up to this many blocks use two kernels

constexpr int kExpandThreads = 256;
constexpr int kExpandTile = 1024;                  // pairs per expansion block

constexpr int kRadixThreads = 256;
constexpr int kRadixRounds = 16;
constexpr int kRadixTile = kRadixThreads * kRadixRounds;  // keys per radix block
constexpr int kRadixBits = 8;

constexpr int kTileThreads = 512;                  // 8 waves per tile block
constexpr int kWaves = kTileThreads / 64;
constexpr int kBatch = 512;                        // pairs staged per LDS batch
constexpr int kSub = 4;                            // sub-brick edge (voxels)
constexpr int kMaxSubTile = 16;                    // largest sub-tile edge the tile kernels render

// -log2(e)/2: q' = kQScale * q, w = exp(-q/2) = exp2(q').
constexpr double kQScaleD = -0.72134752044448170368;  // -0.5 / ln(2)
// q <= 9  <=>  q' >= 9 * kQScale
constexpr float kThr = (float)(9.0 * kQScaleD);

constexpr uint32_t kEmptyBox = 0x80000000u;

thread_local char g_last_detail[256] = "";

int hip_status(hipError_t e, const char* where) {
  if (e == hipSuccess) return GSTV_OK;
  std::snprintf(g_last_detail, sizeof(g_last_detail), "%s: %s", where, hipGetErrorString(e));
  return GSTV_ERR_HIP;
}

#define GSTV_LAUNCH_CHECK(where)                                   \
  do {                                                             \
    int _st = hip_status(hipGetLastError(), where);                \
    if (_st != GSTV_OK) return _st;                                \
  } while (0)

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }
inline int ceil_div_i(int a, int b) { return (a + b - 1) / b; }
inline int64_t ceil_div_l(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Carves consecutive 256-byte aligned regions out of a workspace.
struct Carve {
  char* base;
  size_t off = 0;
  explicit Carve(void* p) : base(static_cast<char*>(p)) {}
  template <typename T>
  T* take(size_t count) {
    off = align_up(off, 256);
    T* p = reinterpret_cast<T*>(base ? base + off : nullptr);
    off += count * sizeof(T);
    return p;
  }
};

// Blocks are dispatched to the 8 XCDs round-robin.  Give each XCD a contiguous
// range of logical block ids so neighbouring tiles (which share Gaussians) hit
// the same L2.  grid is a multiple of 8; the map is a bijection on [0, grid).
__device__ __forceinline__ int xcd_remap(int b, int grid) {
  const int per = grid >> 3;
  return (b & 7) * per + (b >> 3);
}

// A value that is NaN or +-inf (one v_cmp_class).
__device__ __forceinline__ bool non_finite(float x) { return !__builtin_isfinite(x); }

// ---------------------------------------------------------------- scanning --

__device__ __forceinline__ int64_t wave_inclusive_scan(int64_t v, int lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int64_t o = __shfl_up(v, off, 64);
    if (lane >= off) v += o;
  }
  return v;
}

// Exclusive scan across a 256-thread block; *total gets the block sum.
__device__ __forceinline__ int64_t block_exclusive_scan(int64_t v, int64_t* sh, int64_t* total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int64_t inc = wave_inclusive_scan(v, lane);
  if (lane == 63) sh[wid] = inc;
  __syncthreads();
  int64_t base = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kScanThreads / 64; ++i) {
    int64_t s = sh[i];
    base += (i < wid) ? s : 0;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return base + inc - v;
}

// Phase 1: per-block sums of int32 input.
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const int32_t* __restrict__ in, int64_t n,
                                                              int64_t* __restrict__ block_sums) {
  __shared__ int64_t sh[kScanThreads / 64];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  int64_t s = 0;
#pragma unroll
  for (int r = 0; r < kScanItems; ++r) {
    int64_t i = base + (int64_t)r * kScanThreads + threadIdx.x;
    if (i < n) s += in[i];
  }
  int64_t tot;
  block_exclusive_scan(s, sh, &tot);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

// Exclusive scan of the block sums in one block (large inputs only; see
// launch_exclusive_scan); total to *total.
__global__ __launch_bounds__(kScanThreads) void k_scan_block_sums(int64_t* __restrict__ sums, int nb,
                                                                  int64_t* __restrict__ total) {
  __shared__ int64_t sh[kScanThreads / 64];
  int64_t carry = 0;
  for (int base = 0; base < nb; base += kScanThreads) {
    const int i = base + threadIdx.x;
    int64_t v = i < nb ? sums[i] : 0;
    int64_t tot;
    int64_t ex = block_exclusive_scan(v, sh, &tot);
    if (i < nb) sums[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0 && total != nullptr) *total = carry;
}

// Phase 2: exclusive scan inside each block plus the block's offset.  With
// `scanned` the block sums were scanned in place by k_scan_block_sums;
// otherwise each block sums the preceding blocks' sums itself (a few hundred
// values, read from L2: one pass fewer) and the last block writes the grand
// total to *total when given.  Each thread owns kScanItems consecutive items.
// out may alias in.
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const int32_t* in, int64_t n,
                                                             const int64_t* __restrict__ block_sums, bool scanned,
                                                             int64_t* __restrict__ total, int32_t* out) {
  __shared__ int64_t sh[kScanThreads / 64];
  int64_t offset;
  if (scanned) {
    offset = block_sums[blockIdx.x];
  } else {
    int64_t part = 0;
    for (int i = threadIdx.x; i < (int)blockIdx.x; i += kScanThreads) part += block_sums[i];
    block_exclusive_scan(part, sh, &offset);
  }
  const int64_t first = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int32_t v[kScanItems];
  int64_t s = 0;
#pragma unroll
  for (int r = 0; r < kScanItems; ++r) {
    const int64_t i = first + r;
    v[r] = i < n ? in[i] : 0;
    s += v[r];
  }
  int64_t tot;
  int64_t run = block_exclusive_scan(s, sh, &tot) + offset;
#pragma unroll
  for (int r = 0; r < kScanItems; ++r) {
    const int64_t i = first + r;
    if (i < n) out[i] = (int32_t)run;
    run += v[r];
  }
  if (!scanned && total != nullptr && threadIdx.x == 0 && blockIdx.x == gridDim.x - 1) *total = offset + tot;
}

// Device-wide exclusive scan of n int32 values into int32 (values must fit).
// block_sums needs ceil(n / kScanTile) entries.
int launch_exclusive_scan(const int32_t* in, int64_t n, int32_t* out, int64_t* block_sums,
                          int64_t* total, hipStream_t s) {
  const int nb = (int)ceil_div_l(n, kScanTile);
  if (nb == 0) return GSTV_OK;
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kScanThreads), 0, s, in, n, block_sums);
  GSTV_LAUNCH_CHECK("scan_reduce");
  // Up to kScanFoldBlocks blocks, each block sums its predecessors' sums
  // itself (two kernels); beyond, the block sums get a pass of their own.
  const bool scanned = nb > kScanFoldBlocks;
  if (scanned) {
    hipLaunchKernelGGL(k_scan_block_sums, dim3(1), dim3(kScanThreads), 0, s, block_sums, nb, total);
    GSTV_LAUNCH_CHECK("scan_block_sums");
  }
  hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kScanThreads), 0, s, in, n, (const int64_t*)block_sums, scanned,
                     total, out);
  GSTV_LAUNCH_CHECK("scan_apply");
  return GSTV_OK;
}

__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_f32(float* __restrict__ p, int64_t n, float v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ----------------------------------------------------------------- binning --

// Record of one Gaussian for the tile kernels (built by k_prep).  The
// exponent q' = -log2(e)/2 * q of voxel offset u (axes D, H, W = u0, u1, u2)
// is kept as squares of linear forms, eliminated in the kernels' loop order
// (H innermost, then D, W fixed per lane):
//   q' = -(y1^2 + y2^2 + y3w u2^2),  y1 = y1h u1 + y1d u0 + y1w u2,
//                                     y2 = y2d u0 + y2w u2.
struct __align__(16) GRec {
  float c0, c1, c2, rr;      // centre (voxel units, axes D,H,W), density real
  float ri, y1h, y1d, y1w;   // density imag; first linear form
  float y2d, y2w, y3w, e0;   // second linear form, u2^2 coefficient; extent D
  float e1, e2;              // extents H, W
  int32_t gid;               // the Gaussian's own index
  float pad;
};
static_assert(sizeof(GRec) == 64, "record must be one 64-byte line");

// The fields of record i that the tile kernels read (gid only with kGid), as
// three 16-byte loads and one of 12 or 8 bytes.  No loaded word goes unused:
// the compiler may give the register of an unused word to another value while
// its load is in flight, and must then wait for the load at once, which would
// undo reading records one batch ahead.  The tail's offset is hidden from the
// compiler (an empty asm on a register), which would otherwise widen the
// short load to the whole 16 bytes.
template <bool kGid>
__device__ __forceinline__ GRec load_rec(const GRec* __restrict__ rec, int i) {
  const float4* p = reinterpret_cast<const float4*>(rec + i);
  const float4 a = p[0], b = p[1], c = p[2];
  int tail = 48;
  asm volatile("" : "+s"(tail));
  const float* q = reinterpret_cast<const float*>(reinterpret_cast<const char*>(rec + i) + tail);
  GRec r;
  r.c0 = a.x;
  r.c1 = a.y;
  r.c2 = a.z;
  r.rr = a.w;
  r.ri = b.x;
  r.y1h = b.y;
  r.y1d = b.z;
  r.y1w = b.w;
  r.y2d = c.x;
  r.y2w = c.y;
  r.y3w = c.z;
  r.e0 = c.w;
  r.e1 = q[0];
  r.e2 = q[1];
  r.gid = kGid ? __float_as_int(q[2]) : -1;
  r.pad = 0.0f;
  return r;
}

struct BinGeom {
  int size[3];
  int nt[3];
  int tile;
  int tile_shift;   // log2(tile) for a power of two, else -1
  const GRec* rec;  // non-null: pairs only where the records' extent boxes meet the tile
  // Non-null (with rec): per Gaussian, the live tiles of a tile box of at most
  // kLiveMaxTiles tiles, bit k for the box's k-th tile in row-major local order
  // (written by the counting pass, read by the expansion); larger boxes keep
  // every tile.
  uint32_t* live;
};

// Tile boxes of at most this many tiles carry a live-tile mask.
constexpr int kLiveMaxTiles = 32;

// Position of the set bit of rank l (0-based) in m; m has more than l set bits.
__device__ __forceinline__ int nth_set_bit(uint32_t m, int l) {
  int pos = 0;
#pragma unroll
  for (int w = 16; w >= 1; w >>= 1) {
    const int c = __popc(m & ((1u << w) - 1u));
    if (l >= c) {
      l -= c;
      m >>= w;
      pos += w;
    }
  }
  return pos;
}

// Local index l of a Gaussian's pair -> (ld, lh, lw) in its tile box of
// tcnt[1] x tcnt[2] tiles per D slab; with a live mask, l counts live tiles.
__device__ __forceinline__ void box_local(int l, const int tcnt[3], const uint32_t* live, int g, int& ld, int& lh,
                                          int& lw) {
  const int n12 = tcnt[1] * tcnt[2];
  if (live != nullptr && tcnt[0] * n12 <= kLiveMaxTiles) l = nth_set_bit(live[g], l);
  ld = l / n12;
  const int rem = l - ld * n12;
  lh = rem / tcnt[2];
  lw = rem - lh * tcnt[2];
}

inline int tile_shift_of(int tile) {
  int sh = 0;
  while ((1 << sh) < tile) ++sh;
  return (1 << sh) == tile ? sh : -1;
}

// float -> int32 conversion with truncation, saturation and NaN -> 0 (the
// semantics of a device-side float->int cast, which the reference's GPU path
// applies in `.int()` at gaussian/tile_voxelizer.py:98-99).
__device__ __forceinline__ int sat_f2i(float x) {
  if (!(x == x)) return 0;
  if (x >= 2147483648.0f) return INT_MAX;
  if (x <= -2147483648.0f) return INT_MIN;
  return (int)x;
}

// Non-negative x / tile (a shift for power-of-two tiles).
__device__ __forceinline__ int tile_of(int x, const BinGeom& geo) {
  return geo.tile_shift >= 0 ? (x >> geo.tile_shift) : x / geo.tile;
}

// Tile box of one Gaussian; gaussian/tile_voxelizer.py:97-116.  With tc / te
// (centre and extents of the Gaussian's record), the box is cut to the tiles
// that the extent box [ceil(c - e), floor(c + e)] meets: the tile kernels skip
// every other pair of the tile lists (the box test of stage_pair, with the same
// float bounds), so dropping them changes no evaluation (only how a tile's
// pairs fall into the forward's batches); a NaN bound restricts
// nothing, as there.
__device__ __forceinline__ void tile_box_v(const float cref[3], float r, const BinGeom& geo, const float* tc,
                                           const float* te, int tmin[3], int tcnt[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float c = cref[a];
    int lo = sat_f2i(floorf(c - r));
    int hi = sat_f2i(ceilf(c + r));
    lo = min(max(lo, 0), geo.size[a] - 1);
    hi = min(max(hi, 0), geo.size[a] - 1);
    int tl = min(tile_of(lo, geo), geo.nt[a] - 1);
    int th = min(tile_of(hi, geo), geo.nt[a] - 1);
    if (tc != nullptr) {
      float blo = ceilf(tc[a] - te[a]), bhi = floorf(tc[a] + te[a]);
      if (!(blo == blo)) blo = -INFINITY;
      if (!(bhi == bhi)) bhi = INFINITY;
      const float last = (float)(geo.size[a] - 1);
      if (blo > bhi || blo > last || bhi < 0.0f) {
        th = tl - 1;
      } else {
        if (blo > 0.0f) tl = max(tl, tile_of((int)blo, geo));
        if (bhi < last) th = min(th, tile_of((int)bhi, geo));
      }
    }
    tmin[a] = tl;
    tcnt[a] = max(th - tl + 1, 0);  // a negative radius can invert the box: no pairs
  }
}

// tile_box_v of Gaussian g read from center_vox / radius_vox.
__device__ __forceinline__ void tile_box(const float* __restrict__ center, const float* __restrict__ radius, int g,
                                         const BinGeom& geo, const float* tc, const float* te, int tmin[3],
                                         int tcnt[3]) {
  const float cref[3] = {center[3 * (int64_t)g + 0], center[3 * (int64_t)g + 1], center[3 * (int64_t)g + 2]};
  tile_box_v(cref, radius[g], geo, tc, te, tmin, tcnt);
}

// tile_box of Gaussian g, cut to its record's extent box when geo.rec is set.
__device__ __forceinline__ void gauss_tile_box(const float* __restrict__ center, const float* __restrict__ radius,
                                               int g, const BinGeom& geo, int tmin[3], int tcnt[3]) {
  if (geo.rec != nullptr) {
    const GRec& rg = geo.rec[g];
    const float tc[3] = {rg.c0, rg.c1, rg.c2}, te[3] = {rg.e0, rg.e1, rg.e2};
    tile_box(center, radius, g, geo, tc, te, tmin, tcnt);
  } else {
    tile_box(center, radius, g, geo, nullptr, nullptr, tmin, tcnt);
  }
}

// The pair total the binning kernels work to.  With dev null it is cap (the
// host knows the total).  Otherwise the kernels read it from device memory and
// cap is the capacity the caller sized the grids and arrays for: a total above
// it (or negative) reads as 0, so every kernel writes in bounds and the tile
// lists come out empty; the caller compares the total with cap afterwards.
struct PairTotal {
  const int64_t* dev;
  int cap;
};

__device__ __forceinline__ int pair_total(const PairTotal& t) {
  if (t.dev == nullptr) return t.cap;
  const int64_t v = *t.dev;
  return (v >= 0 && v <= (int64_t)t.cap) ? (int)v : 0;
}

// Pairs per Gaussian; counts[n] = 0 so that the scan of n + 1 entries leaves
// the total in its last entry.
__global__ __launch_bounds__(256) void k_bin_count(const float* __restrict__ center,
                                                   const float* __restrict__ radius, int n, BinGeom geo,
                                                   int32_t* __restrict__ counts) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g == n) counts[n] = 0;
  if (g >= n) return;
  int tmin[3], tcnt[3];
  gauss_tile_box(center, radius, g, geo, tmin, tcnt);
  counts[g] = tcnt[0] * tcnt[1] * tcnt[2];  // <= num_tiles < 2^31
}

// Largest i in [0, len) with a[i] <= x (a non-decreasing, a[0] <= x).
__device__ __forceinline__ int last_le(const int32_t* a, int len, int32_t x) {
  int lo = 0, hi = len - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (a[mid] <= x) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Expands pairs in Gaussian-major, row-major-local order
// (gaussian/tile_voxelizer.py:126-145).  Each block owns kExpandTile
// consecutive pair slots; it locates the Gaussians covering them by binary
// search in the offsets, caches their tile boxes in LDS, and writes keys
// (tile id) and values (Gaussian id) coalesced.
__global__ __launch_bounds__(kExpandThreads) void k_bin_expand(
    const float* __restrict__ center, const float* __restrict__ radius, int n, BinGeom geo,
    const int32_t* __restrict__ offsets, PairTotal pt, int32_t* __restrict__ keys,
    int32_t* __restrict__ vals) {
  __shared__ int32_t sh_off[kExpandTile + 1];
  __shared__ int32_t sh_t0[kExpandTile], sh_t1[kExpandTile], sh_t2[kExpandTile];
  __shared__ int32_t sh_n0[kExpandTile], sh_n1[kExpandTile], sh_n2[kExpandTile];
  __shared__ int sh_g0, sh_ng;
  const int total = pair_total(pt);
  const int j0 = blockIdx.x * kExpandTile;
  if (j0 >= total) return;  // the grid covers the capacity
  const int j1 = min(j0 + kExpandTile, total);
  if (threadIdx.x == 0) {
    const int g_first = last_le(offsets, n + 1, j0);
    const int g_last = last_le(offsets, n + 1, j1 - 1);
    sh_g0 = g_first;
    sh_ng = g_last - g_first + 1;
  }
  __syncthreads();
  const int g0 = sh_g0, ng = sh_ng;
  const int ntw = geo.nt[2];
  const int nthw = geo.nt[1] * geo.nt[2];
  if (ng > kExpandTile) {
    // Only reachable when Gaussians with zero pairs sit inside this block's
    // range: locate each pair's Gaussian in global memory instead.
    for (int j = j0 + threadIdx.x; j < j1; j += kExpandThreads) {
      const int g = last_le(offsets, n + 1, j);
      int tmin[3], tcnt[3];
      gauss_tile_box(center, radius, g, geo, tmin, tcnt);
      int ld, lh, lw;
      box_local(j - offsets[g], tcnt, geo.live, g, ld, lh, lw);
      keys[j] = (tmin[0] + ld) * nthw + (tmin[1] + lh) * ntw + (tmin[2] + lw);
      vals[j] = g;
    }
    return;
  }
  for (int i = threadIdx.x; i < ng; i += kExpandThreads) {
    int tmin[3], tcnt[3];
    gauss_tile_box(center, radius, g0 + i, geo, tmin, tcnt);
    sh_off[i] = offsets[g0 + i];
    sh_t0[i] = tmin[0];
    sh_t1[i] = tmin[1];
    sh_t2[i] = tmin[2];
    sh_n0[i] = tcnt[0];
    sh_n1[i] = tcnt[1];
    sh_n2[i] = tcnt[2];
  }
  __syncthreads();
  for (int j = j0 + threadIdx.x; j < j1; j += kExpandThreads) {
    const int i = last_le(sh_off, ng, j);
    const int tcnt[3] = {sh_n0[i], sh_n1[i], sh_n2[i]};
    int ld, lh, lw;
    box_local(j - sh_off[i], tcnt, geo.live, g0 + i, ld, lh, lw);
    keys[j] = (sh_t0[i] + ld) * nthw + (sh_t1[i] + lh) * ntw + (sh_t2[i] + lw);
    vals[j] = g0 + i;
  }
}

// ---- pair generation inside the first radix pass ----
//
// The first radix pass does not read materialised pairs: each block expands
// its kRadixTile consecutive pair slots (Gaussian-major, row-major-local
// order, gaussian/tile_voxelizer.py:126-145) from the Gaussians' tile boxes.
// The block's Gaussians are located by two binary searches in the offsets,
// their boxes go to LDS (first tile id, tiles per D slab, tiles per H row),
// and every slot gets its Gaussian from an LDS slot -> Gaussian map.

// Gaussians whose boxes a generating block keeps in LDS.  A block of
// kRadixTile slots spanning more (only Gaussians with a few pairs each, or none)
// searches per slot instead; the smaller table keeps LDS per block under 40 KB
// so that four blocks share a CU.
constexpr int kGenCap = 1920;

struct GenLds {
  int32_t off[kGenCap];
  int32_t k0[kGenCap];
  int32_t n12[kGenCap];
  int32_t n2[kGenCap];
  uint16_t gidx[kRadixTile];  // 16-byte aligned: follows 4 * kGenCap int32
  uint32_t wmax[kRadixThreads / 64];
  int g0, ng;
};
static_assert((4 * kGenCap * sizeof(int32_t)) % 16 == 0, "gidx must be 16-byte aligned");
// The owner scan gives each thread kRadixTile / kRadixThreads consecutive
// slots, read as 16-byte vectors of 8 slots.
constexpr int kOwnerVecs = kRadixTile / kRadixThreads / 8;
static_assert(kOwnerVecs >= 1 && kRadixTile == kRadixThreads * 8 * kOwnerVecs, "slots per thread: a multiple of 8");

struct ScatterLds {
  int32_t key[kRadixTile];
  int32_t val[kRadixTile];
  int32_t wcnt[kRadixThreads / 64][1 << kRadixBits];
  int32_t local[1 << kRadixBits];
  int32_t base[1 << kRadixBits];
  int64_t scan[kRadixThreads / 64];
};

// Owner Gaussian of the first slot of every kRadixTile-slot block
// (first[nblk] = owner of the last slot), written by the Gaussians themselves:
// no per-block search.  Thread g also presets tile g's range to (INT_MAX, 0)
// for the last radix pass's merge (one launch for both).
__global__ void k_block_first(const int32_t* __restrict__ offsets, int n, PairTotal pt,
                              int32_t* __restrict__ first, int32_t* __restrict__ ranges, int num_tiles) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < num_tiles) {
    ranges[2 * (int64_t)g] = INT_MAX;
    ranges[2 * (int64_t)g + 1] = 0;
  }
  const int total = pair_total(pt);
  if (g >= n || total == 0) return;  // total 0: no block reads first
  const int lo = offsets[g], hi = offsets[g + 1];
  if (lo >= hi) return;
  for (int b = (lo + kRadixTile - 1) / kRadixTile; b * (int64_t)kRadixTile < hi; ++b) first[b] = g;
  if (lo <= total - 1 && total - 1 < hi) first[(total + kRadixTile - 1) / kRadixTile] = g;
}

// Block-wide: maps the slots [j0, j1) of block b to their Gaussians
// (first[b] .. first[b + 1], an upper bound of the last owner).  Returns false
// (uniformly) when more than kGenCap Gaussians span the block, or when the
// volume has more than 2^20 tiles (the bound of gen_pair's divisions);
// callers then search per slot.
__device__ __forceinline__ bool gen_setup(GenLds& L, const float* __restrict__ center, const float* __restrict__ radius,
                          const BinGeom& geo, const int32_t* __restrict__ offsets,
                          const int32_t* __restrict__ first, int j0, int j1) {
  const int b = j0 / kRadixTile;
  const int g0 = first[b];
  const int ng = first[b + 1] - g0 + 1;
  if (threadIdx.x == 0) L.g0 = g0;
  if (ng > kGenCap || geo.nt[0] * geo.nt[1] * geo.nt[2] > (1 << 20)) return false;
  const int ntw = geo.nt[2];
  const int nthw = geo.nt[1] * geo.nt[2];
  uint4* g16 = reinterpret_cast<uint4*>(L.gidx);
  const uint4 zero4 = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
  for (int q = 0; q < kOwnerVecs; ++q) g16[kOwnerVecs * threadIdx.x + q] = zero4;
  for (int i = threadIdx.x; i < ng; i += kRadixThreads) {
    int tmin[3], tcnt[3];
    gauss_tile_box(center, radius, g0 + i, geo, tmin, tcnt);
    L.off[i] = offsets[g0 + i];
    L.k0[i] = tmin[0] * nthw + tmin[1] * ntw + tmin[2];
    const int n12 = tcnt[1] * tcnt[2];
    if (geo.live != nullptr && tcnt[0] * n12 <= kLiveMaxTiles) {
      // A box with a live mask: n12, n2 <= 32 packed as -1 - (n12 << 6 | n2),
      // and the mask in place of n2.
      L.n12[i] = -1 - ((n12 << 6) | tcnt[2]);
      L.n2[i] = (int32_t)geo.live[g0 + i];
    } else {
      L.n12[i] = n12;
      L.n2[i] = tcnt[2];
    }
  }
  __syncthreads();
  // Slot -> Gaussian map.  The owner of a slot is the last Gaussian starting
  // at or before it: each Gaussian with pairs that starts inside the block
  // marks its first slot (Gaussian 0, first[b], owns slot 0 through the zero
  // fill), and an inclusive max-scan over the slots spreads the marks.
  for (int i = threadIdx.x + 1; i < ng; i += kRadixThreads) {
    const int o = L.off[i];
    const int next = i + 1 < ng ? L.off[i + 1] : offsets[g0 + ng];
    if (o < next && o < j1) L.gidx[o - j0] = (uint16_t)i;  // o > j0: slot j0 belongs to Gaussian 0
  }
  __syncthreads();
  // Each thread scans its consecutive slots (two per 32-bit word), then the
  // running maxima of the threads are combined across the wave and the waves.
  uint32_t v[4 * kOwnerVecs];
#pragma unroll
  for (int q = 0; q < kOwnerVecs; ++q) {
    const uint4 a = g16[kOwnerVecs * threadIdx.x + q];
    v[4 * q + 0] = a.x;
    v[4 * q + 1] = a.y;
    v[4 * q + 2] = a.z;
    v[4 * q + 3] = a.w;
  }
  uint32_t run = 0;
#pragma unroll
  for (int k = 0; k < 4 * kOwnerVecs; ++k) {
    const uint32_t lo = max(v[k] & 0xffffu, run);
    const uint32_t hi = max(v[k] >> 16, lo);
    v[k] = lo | (hi << 16);
    run = hi;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t incl = run;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(incl, d, 64);
    if (lane >= d) incl = max(incl, o);
  }
  uint32_t excl = __shfl_up(incl, 1, 64);
  if (lane == 0) excl = 0;
  if (lane == 63) L.wmax[wid] = incl;
  __syncthreads();
  for (int w = 0; w < wid; ++w) excl = max(excl, L.wmax[w]);
#pragma unroll
  for (int k = 0; k < 4 * kOwnerVecs; ++k) {
    // Both halves are at most 0xffff: a per-half max of v[k] and excl.
    const uint32_t lo = max(v[k] & 0xffffu, excl), hi = max(v[k] >> 16, excl);
    v[k] = lo | (hi << 16);
  }
#pragma unroll
  for (int q = 0; q < kOwnerVecs; ++q)
    g16[kOwnerVecs * threadIdx.x + q] = make_uint4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
  __syncthreads();
  return true;
}

// floor(l / d) for 0 <= l < 2^20 and d >= 1 from one float reciprocal.  The
// fraction of (l + 1/2) / d is at least 1/(2d) away from an integer; with
// v_rcp_f32's 1-ulp reciprocal the product's relative error is below
// 1.5 * 2^-23, an absolute error below (l + 1/2) / d * 1.5 * 2^-23 < 1/(2d)
// for l < 2^20 (with a margin of more than 2x), so truncation gives the exact
// quotient.
__device__ __forceinline__ int div_small(int l, int d) {
  return (int)(((float)l + 0.5f) * __builtin_amdgcn_rcpf((float)d));
}

// Key (tile id) and value (Gaussian id) of pair slot j (j0 <= j < j1).  A
// slot's local index l is below its Gaussian's pair count <= num_tiles, which
// gen_setup bounds by 2^20 for div_small.
__device__ __forceinline__ void gen_pair(const GenLds& L, int j, int j0, const BinGeom& geo, int32_t& key,
                                         int32_t& val) {
  const int i = L.gidx[j - j0];
  int l = j - L.off[i];
  int n12 = L.n12[i], n2 = L.n2[i];
  if (n12 < 0) {
    // A live-masked box: the l-th live tile.
    l = nth_set_bit((uint32_t)n2, l);
    const int p = -1 - n12;
    n12 = p >> 6;
    n2 = p & 63;
  }
  const int ld = div_small(l, n12);
  const int rem = l - ld * n12;
  const int lh = div_small(rem, n2);
  const int lw = rem - lh * n2;
  key = L.k0[i] + ld * (geo.nt[1] * geo.nt[2]) + lh * geo.nt[2] + lw;
  val = L.g0 + i;
}

// Pair slot j located with a binary search in the global offsets.
__device__ __forceinline__ void gen_pair_global(const float* __restrict__ center, const float* __restrict__ radius,
                                                int n, const BinGeom& geo, const int32_t* __restrict__ offsets, int j,
                                                int32_t& key, int32_t& val) {
  const int g = last_le(offsets, n + 1, j);
  int tmin[3], tcnt[3];
  gauss_tile_box(center, radius, g, geo, tmin, tcnt);
  int ld, lh, lw;
  box_local(j - offsets[g], tcnt, geo.live, g, ld, lh, lw);
  key = (tmin[0] + ld) * (geo.nt[1] * geo.nt[2]) + (tmin[1] + lh) * geo.nt[2] + (tmin[2] + lw);
  val = g;
}

// Where the pairs of a radix pass come from: materialised arrays, or (first
// pass) generated from the Gaussians' tile boxes.
struct PairSource {
  const int32_t* keys;
  const int32_t* vals;
  const float* center;
  const float* radius;
  const int32_t* offsets;
  const int32_t* first;  // owner of each block's first slot (k_block_first)
  int n;
  BinGeom geo;
};

// Radix upsweep: digit histogram of each kRadixTile-key block, stored
// digit-major (hist[digit * nblocks + block]) so one exclusive scan gives
// every block's scatter base per digit.
template <bool kGen>
__global__ __launch_bounds__(kRadixThreads, 4) void k_radix_hist(PairSource src, PairTotal pt, int shift, int bits,
                                                             int nblocks, int32_t* __restrict__ hist) {
  struct NoGen {};
  __shared__ int32_t sh_h[1 << kRadixBits];
  __shared__ std::conditional_t<kGen, GenLds, NoGen> gen_lds;
  const int total = pair_total(pt);
  const int mask = (1 << bits) - 1;
  const int base = blockIdx.x * kRadixTile;
  if (base >= total) {
    // A block past the total (the grid covers the capacity): empty histogram.
    if (threadIdx.x < (1 << bits)) hist[threadIdx.x * nblocks + blockIdx.x] = 0;
    return;
  }
  sh_h[threadIdx.x] = 0;
  bool gen_ok = false;
  if constexpr (kGen)
    gen_ok = gen_setup(gen_lds, src.center, src.radius, src.geo, src.offsets, src.first, base,
                       min(base + kRadixTile, total));
  __syncthreads();
  if constexpr (!kGen) {
    if (base + kRadixTile <= total) {
      // A full block: 16-byte loads, all in flight before the first atomic.
      const int4* kv = reinterpret_cast<const int4*>(src.keys + base);
      int4 q[kRadixRounds / 4];
#pragma unroll
      for (int r = 0; r < kRadixRounds / 4; ++r) q[r] = kv[r * kRadixThreads + threadIdx.x];
#pragma unroll
      for (int r = 0; r < kRadixRounds / 4; ++r) {
        atomicAdd(&sh_h[(q[r].x >> shift) & mask], 1);
        atomicAdd(&sh_h[(q[r].y >> shift) & mask], 1);
        atomicAdd(&sh_h[(q[r].z >> shift) & mask], 1);
        atomicAdd(&sh_h[(q[r].w >> shift) & mask], 1);
      }
      __syncthreads();
      if (threadIdx.x < (1 << bits)) hist[threadIdx.x * nblocks + blockIdx.x] = sh_h[threadIdx.x];
      return;
    }
  }
#pragma unroll 4
  for (int r = 0; r < kRadixRounds; ++r) {
    const int idx = base + r * kRadixThreads + threadIdx.x;
    if (idx < total) {
      int32_t key, val;
      if constexpr (!kGen) {
        key = src.keys[idx];
      } else {
        if (gen_ok)
          gen_pair(gen_lds, idx, base, src.geo, key, val);
        else
          gen_pair_global(src.center, src.radius, src.n, src.geo, src.offsets, idx, key, val);
      }
      atomicAdd(&sh_h[(key >> shift) & mask], 1);
    }
  }
  __syncthreads();
  if (threadIdx.x < (1 << bits)) hist[threadIdx.x * nblocks + blockIdx.x] = sh_h[threadIdx.x];
}

// Radix downsweep: stable scatter with coalesced writes.  Each wave owns a
// contiguous quarter of the block's keys and ranks equal digits with a ballot
// multisplit against a per-wave running count (no block barriers per round);
// the per-digit wave prefixes then give every key its position in a
// block-local, digit-sorted copy in LDS, which is written out so that
// consecutive threads store consecutive addresses of a digit's run.
//
// kLast (final pass): only the values are written (they are pair_ids), and
// tile_ranges come out of the same pass: the block's copy is sorted by the
// whole key, so the first and last copy of each key in it mark the ends of a
// sub-run of that tile's global range, merged with atomicMin / atomicMax
// into ranges pre-set to (INT_MAX, 0).
template <bool kGen, bool kLast>
__device__ __forceinline__ void radix_scatter(const PairSource& src, const PairTotal& pt, int shift, int bits,
                                              int nblocks,
                                              const int32_t* __restrict__ hist_scanned,
                                              int32_t* __restrict__ keys_out, int32_t* __restrict__ vals_out,
                                              int32_t* __restrict__ ranges) {
  constexpr int kW = kRadixThreads / 64;
  constexpr int kPerWave = kRadixTile / kW;
  constexpr int kRounds = kPerWave / 64;
  union GenScatterLds {
    GenLds gen;
    ScatterLds sc;
  };
  struct ScatterOnly {
    ScatterLds sc;
  };
  __shared__ std::conditional_t<kGen, GenScatterLds, ScatterOnly> L;
  const int total = pair_total(pt);
  if ((int)blockIdx.x * kRadixTile >= total) return;  // past the total: nothing to move
  const int tid = threadIdx.x;
  const int lane = tid & 63, wid = tid >> 6;
  const int ndig = 1 << bits;
  const int mask = ndig - 1;
  const int block0 = blockIdx.x * kRadixTile;
  const int wbase = block0 + wid * kPerWave;

  int32_t kk[kRounds], vv[kRounds], rk[kRounds];
  if constexpr (kGen) {
    const bool ok = gen_setup(L.gen, src.center, src.radius, src.geo, src.offsets, src.first, block0,
                              min(block0 + kRadixTile, total));
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
      const int idx = wbase + r * 64 + lane;
      kk[r] = 0;
      vv[r] = 0;
      if (idx < total) {
        if (ok)
          gen_pair(L.gen, idx, block0, src.geo, kk[r], vv[r]);
        else
          gen_pair_global(src.center, src.radius, src.n, src.geo, src.offsets, idx, kk[r], vv[r]);
      }
    }
    __syncthreads();  // the LDS is reused below
  } else {
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
      const int idx = wbase + r * 64 + lane;
      kk[r] = idx < total ? src.keys[idx] : 0;
      vv[r] = idx < total ? src.vals[idx] : 0;
    }
  }
  ScatterLds& sh = L.sc;
#pragma unroll
  for (int w = 0; w < kW; ++w) sh.wcnt[w][tid] = 0;
  sh.base[tid] = tid < ndig ? hist_scanned[tid * nblocks + blockIdx.x] : 0;
  __syncthreads();

  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
  for (int r = 0; r < kRounds; ++r) {
    const int idx = wbase + r * 64 + lane;
    const bool valid = idx < total;
    const int dig = (kk[r] >> shift) & mask;
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < bits; ++b) {
      const bool bit = (dig >> b) & 1;
      const uint64_t bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    const int wrank = __popcll(peers & lt_mask);
    const int prev = sh.wcnt[wid][dig];  // read by the whole wave before its leader updates it
    if (valid && wrank == 0) sh.wcnt[wid][dig] = prev + __popcll(peers);
    rk[r] = valid ? prev + wrank : -1;
  }
  __syncthreads();
  {
    // Per digit (thread tid = digit): wave prefixes and the block total, then
    // the block-local exclusive offsets across digits.
    int run = 0;
#pragma unroll
    for (int w = 0; w < kW; ++w) {
      const int c = sh.wcnt[w][tid];
      sh.wcnt[w][tid] = run;
      run += c;
    }
    int64_t tot;
    sh.local[tid] = (int32_t)block_exclusive_scan(run, sh.scan, &tot);
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kRounds; ++r) {
    if (rk[r] >= 0) {
      const int dig = (kk[r] >> shift) & mask;
      const int lpos = sh.local[dig] + sh.wcnt[wid][dig] + rk[r];
      sh.key[lpos] = kk[r];
      sh.val[lpos] = vv[r];
    }
  }
  __syncthreads();
  const int nblk = min(kRadixTile, total - block0);
#pragma unroll 4
  for (int i = tid; i < nblk; i += kRadixThreads) {
    const int32_t key = sh.key[i];
    const int dig = (key >> shift) & mask;
    const int out = sh.base[dig] + (i - sh.local[dig]);
    vals_out[out] = sh.val[i];
    if (!kLast) {
      keys_out[out] = key;
    } else {
      if (i == 0 || sh.key[i - 1] != key) atomicMin(&ranges[2 * (int64_t)key], out);
      if (i == nblk - 1 || sh.key[i + 1] != key) atomicMax(&ranges[2 * (int64_t)key + 1], out + 1);
    }
  }
}

// First pass (pairs generated in the kernel; under 40 KB of LDS, 4 blocks per CU).
template <bool kLast>
__global__ __launch_bounds__(kRadixThreads, 4) void k_radix_scatter_gen(PairSource src, PairTotal pt, int shift, int bits,
                                                                    int nblocks,
                                                                    const int32_t* __restrict__ hist_scanned,
                                                                    int32_t* __restrict__ keys_out,
                                                                    int32_t* __restrict__ vals_out,
                                                                    int32_t* __restrict__ ranges) {
  radix_scatter<true, kLast>(src, pt, shift, bits, nblocks, hist_scanned, keys_out, vals_out, ranges);
}

// Later passes (pairs read from the previous pass; 4 blocks per CU).
template <bool kLast>
__global__ __launch_bounds__(kRadixThreads, 4) void k_radix_scatter(PairSource src, PairTotal pt, int shift, int bits,
                                                                    int nblocks,
                                                                    const int32_t* __restrict__ hist_scanned,
                                                                    int32_t* __restrict__ keys_out,
                                                                    int32_t* __restrict__ vals_out,
                                                                    int32_t* __restrict__ ranges) {
  radix_scatter<false, kLast>(src, pt, shift, bits, nblocks, hist_scanned, keys_out, vals_out, ranges);
}

// tile_ranges of the fused final pass (preset to (INT_MAX, 0) by
// k_block_first): tiles with no pair get (0, 0) as in the reference.
__global__ void k_ranges_fix(int32_t* __restrict__ ranges, int num_tiles) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < num_tiles && ranges[2 * (int64_t)t] == INT_MAX) ranges[2 * (int64_t)t] = 0;
}

// tile_ranges from sorted keys; gaussian/tile_voxelizer.py:151-161.
__global__ void k_ranges(const int32_t* __restrict__ keys, PairTotal pt, int32_t* __restrict__ ranges) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int total = pair_total(pt);
  if (j >= total) return;
  const int t = keys[j];
  if (j == 0 || keys[j - 1] != t) ranges[2 * t] = j;
  if (j == total - 1 || keys[j + 1] != t) ranges[2 * t + 1] = j + 1;
}

// -------------------------------------------------------- Gaussian records --

struct ZAxis {
  float al, ga, ta, rho;
};
struct ZCoefs {
  ZAxis h, w;  // projections on (D, H) and (D, W)
  bool ok;     // false: no bound (extents infinite, or the form is not definite)
};

// The two projections of one Gaussian, built once per call by k_prep (32 B);
// h.rho < 0 stands for "no bound".
struct __align__(16) ZRec {
  ZAxis h, w;
};
static_assert(sizeof(ZRec) == 32, "column-bound record must be 32 bytes");

__device__ __forceinline__ ZCoefs z_load(const ZRec* __restrict__ zt, int g) {
  const ZRec z = zt[g];
  return ZCoefs{z.h, z.w, z.h.rho > 0.0f};
}



// ------------------------------------------------ D extent over a column --
//
// The D rows of a voxel column (H in [y0, y1], W in [x0, x1]) that can hold a
// voxel of the widened ellipsoid E = {u : u^T A u <= R^2} (R = 3.06, the
// radius of the extents e_a = R sqrt(C_aa) + 0.01, C = A^-1) lie in the
// intersection of two intervals: the D extent of E cut by the H strip alone
// and by the W strip alone.  Cutting by the strip u_i in [lo, hi] is the same
// as cutting E's projection on (u_0, u_i), an ellipse with matrix C_sub^-1;
// along u_i = t that ellipse spans
//     u_0 in al t -/+ ga sqrt(rho^2 - t^2),
//     al = C_0i / C_ii,  ga = sqrt(C_00 C_ii - C_0i^2) / C_ii,  rho = R sqrt(C_ii),
// a concave upper and a convex lower boundary whose extremes lie at
// t = +/- ta, ta = al rho^2 / rho_0.  Over [lo, hi] the maximum of u_0 is
// therefore at clamp(ta, lo, hi) and the minimum at clamp(-ta, lo, hi).
// With cofactors of the scaled form A' = kq A (C_ii = cof_ii / det and, by
// Jacobi's identity, C_00 C_ii - C_0i^2 = A_jj / det A):
//     al = cof_0i / cof_ii,  ga = sqrt(A'_jj det A') / cof_ii.
// The determinant cancels badly for needle-shaped Gaussians, so cofactors and
// determinant are formed in fp64 from the fp32 form the kernels evaluate.
// D extent [zmin, zmax] (relative to the centre) of the projected ellipse over
// the strip u_i in [lo, hi].  Clamps are single med3 instructions: with a
// non-finite strip (non-finite centre) the result is arbitrary, which is
// harmless -- such a Gaussian has q = NaN or -inf everywhere and contributes
// nothing.
__device__ __forceinline__ void z_span(const ZAxis& a, float lo, float hi, float& zmin, float& zmax) {
  const float tp = __builtin_amdgcn_fmed3f(a.ta, lo, hi), tm = __builtin_amdgcn_fmed3f(-a.ta, lo, hi);
  const float rr = a.rho * a.rho;
  const float sp = __builtin_amdgcn_sqrtf(__builtin_amdgcn_fmed3f(fmaf(-tp, tp, rr), 0.0f, INFINITY));
  const float sm = __builtin_amdgcn_sqrtf(__builtin_amdgcn_fmed3f(fmaf(-tm, tm, rr), 0.0f, INFINITY));
  zmax = fmaf(a.ga, sp, a.al * tp);
  zmin = fmaf(-a.ga, sm, a.al * tm);
}

// Absolute D rows [zl, zh] of a column from the strip extents, clipped to the
// Gaussian's box rows [lo0, hi0]; a NaN bound falls back to the box.
__device__ __forceinline__ void column_rows(float c0, float hmin, float hmax, float wmin, float wmax, float lo0,
                                            float hi0, float& zl, float& zh) {
  zl = fmaxf(ceilf(c0 + fmaxf(hmin, wmin) - 0.01f), lo0);
  zh = fminf(floorf(c0 + fminf(hmax, wmax) + 0.01f), hi0);
}

// Live tiles of one Gaussian's tile box (tile_box_v) for the binning, bit k
// for the box's k-th tile in row-major local order.  A tile is dead when its
// voxels inside the Gaussian's extent box cannot meet the widened ellipsoid
// that the column bound keeps (R = 3.06): the D rows that the (D, H) and
// (D, W) projections leave over the tile's whole H and W strips are empty, or
// the H rows that the (H, W) projection leaves over its W strip.  Over the
// whole strips each bound contains the tile kernels' per-column bounds, and
// the margin (0.05 voxel against the kernels' 0.01) covers fp32 rounding, so a
// dead tile is one where the tile kernels evaluate no voxel whose q passes the
// cutoff: dropping its pair changes no contribution.  A Gaussian without a
// bound keeps every tile.
constexpr float kLiveMargin = 0.05f;
__device__ __forceinline__ uint32_t live_tiles(const GRec& r, const ZRec& z, const ZAxis& zhw, const BinGeom& geo,
                                               const int tmin[3], const int tcnt[3]) {
  const int cnt = tcnt[0] * tcnt[1] * tcnt[2];
  const uint32_t all = cnt >= 32 ? 0xffffffffu : ((1u << cnt) - 1u);
  // More than 4 tiles along an axis, or no bound: keep the box.
  if (tcnt[0] > 4 || tcnt[1] > 4 || tcnt[2] > 4 || !(z.h.rho > 0.0f) || !(zhw.rho > 0.0f)) return all;
  const float c[3] = {r.c0, r.c1, r.c2}, e[3] = {r.e0, r.e1, r.e2};
  // Voxel range of the box's tiles per axis (the box cut to the extent box).
  float vlo[3][4], vhi[3][4];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float blo = ceilf(c[a] - e[a]), bhi = floorf(c[a] + e[a]);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int t = tmin[a] + k;
      vlo[a][k] = fmaxf(blo, (float)(t * geo.tile));
      vhi[a][k] = fminf(bhi, (float)(min(t * geo.tile + geo.tile, geo.size[a]) - 1));
    }
  }
  // Spans per H tile (the (D, H) projection) and per W tile (the (D, W) and
  // (H, W) projections): the tests of a tile combine its H and W spans.
  float hmn[4], hmx[4], wmn[4], wmx[4], ylo[4], yhi[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    z_span(z.h, vlo[1][k] - c[1], vhi[1][k] - c[1], hmn[k], hmx[k]);
    z_span(z.w, vlo[2][k] - c[2], vhi[2][k] - c[2], wmn[k], wmx[k]);
    float ymn, ymx;
    z_span(zhw, vlo[2][k] - c[2], vhi[2][k] - c[2], ymn, ymx);
    ylo[k] = ceilf(c[1] + ymn - kLiveMargin);
    yhi[k] = floorf(c[1] + ymx + kLiveMargin);
  }
  uint32_t mask = 0;
#pragma unroll
  for (int i1 = 0; i1 < 4; ++i1) {
#pragma unroll
    for (int i2 = 0; i2 < 4; ++i2) {
      const float zl = ceilf(c[0] + fmaxf(hmn[i1], wmn[i2]) - kLiveMargin);
      const float zh = floorf(c[0] + fminf(hmx[i1], wmx[i2]) + kLiveMargin);
      // Written so that a NaN keeps the tile.
      const bool hdead = fmaxf(ylo[i2], vlo[1][i1]) > fminf(yhi[i2], vhi[1][i1]);
#pragma unroll
      for (int i0 = 0; i0 < 4; ++i0) {
        const bool dead = hdead || (fmaxf(zl, vlo[0][i0]) > fminf(zh, vhi[0][i0]));
        if (i0 < tcnt[0] && i1 < tcnt[1] && i2 < tcnt[2] && !dead)
          mask |= 1u << ((i0 * tcnt[1] + i1) * tcnt[2] + i2);
      }
    }
  }
  return mask;
}

// Record of one Gaussian.  The quadratic form of the reference,
// q = d^T M d with d = (x - c) / hs, is q = u^T A u in voxel units u = x - c
// with A = sym(M) / (hs hs^T).  The record keeps q' = -log2(e)/2 q (so that
// w = exp2(q')) as squares of linear forms -- a Cholesky factorisation of
// P = log2(e)/2 A in the order (u1, u0, u2) -- whose fp32 evaluation stays
// accurate for strongly anisotropic Gaussians: an expanded sum
// sum_ij A_ij u_i u_j cancels terms as large as the condition number times q.
// The factors are formed in fp64 from the fp32 M and rounded once.  A first
// pivot P11 is a sum of squares (never negative); a second pivot that is not
// positive (a form that is not positive definite, from near-degenerate fp32
// M) is clamped to zero, its cross term dropped.  A non-finite M gives NaN
// factors (q = NaN: no voxel passes, as in the reference).
//
// The extent e_a bounds |u_a| on the ellipsoid q <= 9: e_a = 3 hs_a
// sqrt((S^-1)_aa), S = sym(M), evaluated in fp64 and widened by 2% + 0.01
// voxel so that no voxel whose fp32 q passes the cutoff is culled.  A matrix
// that is not positive definite, or not finite, gets an infinite extent (no
// culling).  The column-bound projections (ZRec, below) come from the same
// fp64 form.  Gaussian g's density is rho_r[g * rho_stride],
// rho_i[g * rho_stride] (stride 2 with rho_i = rho_r + 1 reads an interleaved
// complex64 array).
//
// counts (optional): the Gaussian's pairs for binning cut to its extent box
// (tile_box with this record; radius and geo as for k_bin_count), and
// counts[n] = 0 for the scan.
// Record, column-bound record (and, with counts, pair count) of Gaussian g
// from its centre c and matrix mm (the fp32 values of center_vox and cov_inv).
__device__ __forceinline__ void prep_one(int g, const float c[3], const float mm[9], float hs0, float hs1, float hs2,
                                         const float* __restrict__ rho_r, const float* __restrict__ rho_i,
                                         int rho_stride, GRec* __restrict__ rec, ZRec* __restrict__ zt, float rad,
                                         const BinGeom& geo, int32_t* __restrict__ counts) {
  GRec r;
  r.c0 = c[0];
  r.c1 = c[1];
  r.c2 = c[2];
  r.rr = rho_r[(int64_t)g * rho_stride];
  r.ri = rho_i[(int64_t)g * rho_stride];
  // Voxel-unit form, scaled by log2(e)/2 (P) for the factors.
  const double h0 = hs0, h1 = hs1, h2 = hs2;
  const double s00 = mm[0], s11 = mm[4], s22 = mm[8];
  const double s01 = 0.5 * ((double)mm[1] + mm[3]);
  const double s02 = 0.5 * ((double)mm[2] + mm[6]);
  const double s12 = 0.5 * ((double)mm[5] + mm[7]);
  const double A00 = s00 / (h0 * h0), A11 = s11 / (h1 * h1), A22 = s22 / (h2 * h2);
  const double A01 = s01 / (h0 * h1), A02 = s02 / (h0 * h2), A12 = s12 / (h1 * h2);
  {
    const double k = -kQScaleD;
    const double p11 = k * A11, p01 = k * A01, p12 = k * A12, p00 = k * A00, p02 = k * A02, p22 = k * A22;
    const double y1h = p11 > 0.0 ? sqrt(p11) : 0.0;
    const double y1d = y1h > 0.0 ? p01 / y1h : 0.0, y1w = y1h > 0.0 ? p12 / y1h : 0.0;
    const double q00 = p00 - y1d * y1d, q02 = p02 - y1d * y1w, q22 = p22 - y1w * y1w;
    const double y2d = q00 > 0.0 ? sqrt(q00) : 0.0;
    const double y2w = y2d > 0.0 ? q02 / y2d : 0.0;
    const double y3w = q22 - y2w * y2w;
    const bool fin = isfinite(p00) && isfinite(p11) && isfinite(p22) && isfinite(p01) && isfinite(p02) &&
                     isfinite(p12);
    r.y1h = fin ? (float)y1h : NAN;
    r.y1d = fin ? (float)y1d : NAN;
    r.y1w = fin ? (float)y1w : NAN;
    r.y2d = fin ? (float)y2d : NAN;
    r.y2w = fin ? (float)y2w : NAN;
    r.y3w = fin ? (float)y3w : NAN;
  }
  // Extents in fp64.
  const double cof00 = s11 * s22 - s12 * s12;
  const double cof11 = s00 * s22 - s02 * s02;
  const double cof22 = s00 * s11 - s01 * s01;
  const double cof01 = s02 * s12 - s01 * s22;
  const double cof02 = s01 * s12 - s02 * s11;
  const double det = s00 * cof00 + s01 * cof01 + s02 * cof02;
  float e0 = INFINITY, e1 = INFINITY, e2 = INFINITY;
  const bool pd = (s00 > 0.0) && (cof22 > 0.0) && (det > 0.0) && isfinite(det) && cof00 > 0.0 &&
                  cof11 > 0.0;
  if (pd) {
    const double f = 3.0 * 1.02;
    const double x0 = f * (double)hs0 * sqrt(cof00 / det) + 0.01;
    const double x1 = f * (double)hs1 * sqrt(cof11 / det) + 0.01;
    const double x2 = f * (double)hs2 * sqrt(cof22 / det) + 0.01;
    if (isfinite(x0) && isfinite(x1) && isfinite(x2)) {
      e0 = (float)x0;
      e1 = (float)x1;
      e2 = (float)x2;
    }
  }
  // A non-finite centre gives q = NaN or +inf at every voxel (the form is
  // positive semi-definite), so the reference's cutoff test skips all of
  // them: no contribution and zero gradients.  Such a record gets a finite
  // centre and inverted extents, which every box test reads as empty, so no
  // kernel evaluates it (voxel offsets from a non-finite centre would turn
  // the adjoint's zero-weight moments into NaN).
  if (!(isfinite(r.c0) && isfinite(r.c1) && isfinite(r.c2))) {
    r.c0 = r.c1 = r.c2 = 0.0f;
    e0 = e1 = e2 = -1.0f;
  }
  r.e0 = e0;
  r.e1 = e1;
  r.e2 = e2;
  r.gid = g;
  r.pad = 0.0f;
  rec[g] = r;
  // Column-bound projections in voxel units (C = A^-1, cofactors of A):
  //   al = C_0i / C_ii = cof_0i / cof_ii,  ga = sqrt(C_00 C_ii - C_0i^2) / C_ii
  //   = sqrt(A_jj det A) / cof_ii (Jacobi), rho_i = e_i - 0.01 = R sqrt(C_ii),
  //   ta = al rho_i^2 / rho_0; for i = H (j = W) and i = W (j = H).  For the
  // binning's tile test, also the projection on (H, W): the H range along
  // u_2 = t, al = cof_12 / cof_22, ga = sqrt(A_00 det A) / cof_22,
  // ta = al rho_2^2 / rho_1.
  ZRec z;
  ZAxis zhw;
  {
    const double c00 = A11 * A22 - A12 * A12, c11 = A00 * A22 - A02 * A02, c22 = A00 * A11 - A01 * A01;
    const double c01 = A02 * A12 - A01 * A22, c02 = A01 * A12 - A02 * A11, c12 = A01 * A02 - A00 * A12;
    const double dA = A00 * c00 + A01 * c01 + A02 * c02;
    const double r0 = (double)(e0 - 0.01f), r1 = (double)(e1 - 0.01f), r2 = (double)(e2 - 0.01f);
    zhw.al = (float)(c12 / c22);
    zhw.ga = (float)(sqrt(A00 * dA) / c22);
    zhw.rho = (float)r2;
    zhw.ta = (float)(c12 / c22 * (r2 * r2 / r1));
    z.h.al = (float)(c01 / c11);
    z.h.ga = (float)(sqrt(A22 * dA) / c11);
    z.h.rho = (float)r1;
    z.h.ta = (float)(c01 / c11 * (r1 * r1 / r0));
    z.w.al = (float)(c02 / c22);
    z.w.ga = (float)(sqrt(A11 * dA) / c22);
    z.w.rho = (float)r2;
    z.w.ta = (float)(c02 / c22 * (r2 * r2 / r0));
    const bool ok = c11 > 0.0 && c22 > 0.0 && dA > 0.0 && r0 > 0.0 && r1 > 0.0 && r2 > 0.0 && isfinite(r0) &&
                    isfinite(r1) && isfinite(r2) && isfinite(z.h.al) && isfinite(z.h.ga) && isfinite(z.h.ta) &&
                    isfinite(z.w.al) && isfinite(z.w.ga) && isfinite(z.w.ta);
    if (!ok) z.h.rho = z.w.rho = -1.0f;
    // The (H, W) projection bounds nothing unless it, too, is finite.
    if (!(ok && isfinite(zhw.al) && isfinite(zhw.ga) && isfinite(zhw.ta) && zhw.rho > 0.0f)) zhw.rho = -1.0f;
    zt[g] = z;
  }
  if (counts != nullptr) {
    const float tc[3] = {r.c0, r.c1, r.c2}, te[3] = {e0, e1, e2};
    int tmin[3], tcnt[3];
    tile_box_v(c, rad, geo, tc, te, tmin, tcnt);
    int cnt = tcnt[0] * tcnt[1] * tcnt[2];
    if (geo.live != nullptr) {
      uint32_t mask = 0;
      if (cnt > 0 && cnt <= kLiveMaxTiles) {
        mask = live_tiles(r, z, zhw, geo, tmin, tcnt);
        cnt = __popc(mask);
      }
      geo.live[g] = mask;
    }
    counts[g] = cnt;
  }
}

__global__ __launch_bounds__(256) void k_prep(const float* __restrict__ center, const float* __restrict__ cov,
                                              const float* __restrict__ rho_r, const float* __restrict__ rho_i,
                                              int rho_stride, const float* __restrict__ half_shape, int n,
                                              GRec* __restrict__ rec, ZRec* __restrict__ zt,
                                              const float* __restrict__ radius, BinGeom geo,
                                              int32_t* __restrict__ counts) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (counts != nullptr && g == n) counts[n] = 0;
  if (g >= n) return;
  const float* m = cov + 9 * (int64_t)g;
  float mm[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) mm[i] = m[i];
  const float c[3] = {center[3 * (int64_t)g + 0], center[3 * (int64_t)g + 1], center[3 * (int64_t)g + 2]};
  prep_one(g, c, mm, half_shape[0], half_shape[1], half_shape[2], rho_r, rho_i, rho_stride, rec, zt,
           counts != nullptr ? radius[g] : 0.0f, geo, counts);
}

// ------------------------------------------------------------ tile kernels --

// A tile of more than 16 voxels per axis is rendered as parts^3 sub-tiles of
// at most 16 (edge `sub` = ceil(tile / parts), parts = ceil(tile / 16)); each
// sub-tile reads its tile's whole pair list and skips the pairs whose box
// misses it, so every voxel sees the same pairs in the same order as with one
// block per tile.  Tiles of at most 16 are their own single sub-tile.
struct VoxGeom {
  int D, H, W, tile;
  int nTD, nTH, nTW, num_tiles;
  int sub;            // sub-tile edge (voxels), <= 16
  int parts;          // sub-tiles per tile axis
  int nSD, nSH, nSW;  // sub-tile grid, nT * parts per axis
  int num_sub;        // sub-tiles
  int nsb;            // sub-bricks per sub-tile axis, ceil(sub / 4)
  int chunks;         // blocks per sub-tile, ceil(nsb^3 / 8)
  int blocks;         // num_sub * chunks
};

// k_forward stages up to kFwdBatch pairs at a time (two per thread); the
// counting kernel and the non-finite pass stage kBatch.
constexpr int kFwdBatch = 2 * kBatch;

struct Stage {
  float c0[kFwdBatch], c1[kFwdBatch], c2[kFwdBatch], rr[kFwdBatch], ri[kFwdBatch];
  float y1h[kFwdBatch], y1d[kFwdBatch], y1w[kFwdBatch], y2d[kFwdBatch], y2w[kFwdBatch], y3w[kFwdBatch];
  int32_t gid[kFwdBatch];  // Gaussian id of each staged pair
  uint32_t box[kFwdBatch];
  uint8_t sbm[kFwdBatch];  // sub-bricks touched, bit 4*j0 + 2*j1 + j2 (tiles of at most 2x2x2 sub-bricks)
  uint16_t list[kWaves][kFwdBatch];
  uint16_t touch[kFwdBatch];  // pairs whose box touches the tile (any order)
  int n_touch;
  // A staged pair touching the tile had a non-finite density (its staged
  // density was zeroed; k_forward adds it after the batches).  Set during
  // staging, never cleared within a tile.
  int any_bad;
  int n_bad;  // pairs listed by the non-finite pass of the current batch
};

struct BlockCtx {
  int tile;
  int o0, o1, o2;  // tile origin (voxels)
  int e0, e1, e2;  // valid tile extent per axis
  int s0, s1, s2;  // this wave's sub-brick origin, tile-local
  int n0, n1, n2;  // valid voxels of the sub-brick per axis (<= 4)
  bool valid;      // the block maps to a tile
  bool active;     // the wave owns at least one valid voxel
  int p_begin, p_end;
};

__device__ __forceinline__ BlockCtx make_ctx(int logical, const VoxGeom& geo, const int32_t* __restrict__ ranges,
                                             int wave) {
  BlockCtx b;
  const int st = logical / geo.chunks;  // sub-tile
  const int chunk = logical - st * geo.chunks;
  const int nshw = geo.nSH * geo.nSW;
  const int sd = st / nshw;
  const int sh = (st - sd * nshw) / geo.nSW;
  const int sw = st - sd * nshw - sh * geo.nSW;
  int td = sd, th = sh, tw = sw, jd = 0, jh = 0, jw = 0;
  if (geo.parts > 1) {
    td = sd / geo.parts;
    th = sh / geo.parts;
    tw = sw / geo.parts;
    jd = sd - td * geo.parts;
    jh = sh - th * geo.parts;
    jw = sw - tw * geo.parts;
  }
  const int tile = (td * geo.nTH + th) * geo.nTW + tw;
  b.tile = tile;
  b.o0 = td * geo.tile + jd * geo.sub;
  b.o1 = th * geo.tile + jh * geo.sub;
  b.o2 = tw * geo.tile + jw * geo.sub;
  // Extent: the sub-tile, cut by its tile's end and the volume's.
  b.e0 = min(min(geo.sub, geo.tile - jd * geo.sub), geo.D - b.o0);
  b.e1 = min(min(geo.sub, geo.tile - jh * geo.sub), geo.H - b.o1);
  b.e2 = min(min(geo.sub, geo.tile - jw * geo.sub), geo.W - b.o2);
  b.valid = st < geo.num_sub && b.e0 > 0 && b.e1 > 0 && b.e2 > 0;
  const int sb = chunk * kWaves + wave;
  const int nsb = geo.nsb;
  const int sb0 = sb / (nsb * nsb);
  const int sb1 = (sb / nsb) % nsb;
  const int sb2 = sb % nsb;
  b.s0 = sb0 * kSub;
  b.s1 = sb1 * kSub;
  b.s2 = sb2 * kSub;
  b.n0 = min(kSub, b.e0 - b.s0);
  b.n1 = min(kSub, b.e1 - b.s1);
  b.n2 = min(kSub, b.e2 - b.s2);
  b.active = (sb < nsb * nsb * nsb) && b.n0 > 0 && b.n1 > 0 && b.n2 > 0;
  b.p_begin = b.valid ? ranges[2 * tile] : 0;
  b.p_end = b.valid ? ranges[2 * tile + 1] : 0;
  return b;
}

// Stages one pair: copies the record fields and computes the Gaussian's
// exact extent clipped to the tile, in tile-local voxel indices, packed as
// 6 x 5 bits (lo0, lo1, lo2, hi0, hi1, hi2); kEmptyBox if it misses the tile.
// Returns whether the box touches the tile.
__device__ __forceinline__ bool stage_pair(Stage& sh, int k, const GRec& r, int gid, const BlockCtx& b) {
  sh.c0[k] = r.c0;
  sh.c1[k] = r.c1;
  sh.c2[k] = r.c2;
  sh.rr[k] = r.rr;
  sh.ri[k] = r.ri;
  sh.y1h[k] = r.y1h;
  sh.y1d[k] = r.y1d;
  sh.y1w[k] = r.y1w;
  sh.y2d[k] = r.y2d;
  sh.y2w[k] = r.y2w;
  sh.y3w[k] = r.y3w;
  sh.gid[k] = gid;
  // fmaxf / fminf return the non-NaN operand, so a NaN centre keeps the whole
  // tile (its q is NaN and fails the cutoff, as in the reference).
  const float lo0 = fmaxf(ceilf(r.c0 - r.e0), (float)b.o0);
  const float lo1 = fmaxf(ceilf(r.c1 - r.e1), (float)b.o1);
  const float lo2 = fmaxf(ceilf(r.c2 - r.e2), (float)b.o2);
  const float hi0 = fminf(floorf(r.c0 + r.e0), (float)(b.o0 + b.e0 - 1));
  const float hi1 = fminf(floorf(r.c1 + r.e1), (float)(b.o1 + b.e1 - 1));
  const float hi2 = fminf(floorf(r.c2 + r.e2), (float)(b.o2 + b.e2 - 1));
  uint32_t box = kEmptyBox;
  if (lo0 <= hi0 && lo1 <= hi1 && lo2 <= hi2) {
    const uint32_t l0 = (uint32_t)((int)lo0 - b.o0), l1 = (uint32_t)((int)lo1 - b.o1),
                   l2 = (uint32_t)((int)lo2 - b.o2);
    const uint32_t h0 = (uint32_t)((int)hi0 - b.o0), h1 = (uint32_t)((int)hi1 - b.o1),
                   h2 = (uint32_t)((int)hi2 - b.o2);
    box = l0 | (l1 << 5) | (l2 << 10) | (h0 << 15) | (h1 << 20) | (h2 << 25);
  }
  sh.sbm[k] = 0;
  sh.box[k] = box;
  return box != kEmptyBox;
}

// Column word of a pair in a tile of at most 8 voxels per axis: for column
// c = 2 j1 + j2 (H block j1, W block j2) the 7-bit field at bit 7c holds
// live (bit 0), first D row (bits 1-3) and rows - 1 (bits 4-6), tile-local;
// 0 for a column the bound leaves empty.
__device__ __forceinline__ uint32_t col_field(int z_lo, int z_hi) {
  return 1u | ((uint32_t)z_lo << 1) | ((uint32_t)(z_hi - z_lo) << 4);
}

// 2x2x2 mask of the sub-bricks that can hold a voxel of the ellipsoid, for a
// pair (centre c, packed box bx) whose box touches a tile of at most 2
// sub-brick per axis: the box's (H, W) column blocks, each with the D rows its
// column bound leaves.  The same rows are returned as a column word in `cols`.
__device__ __forceinline__ uint32_t pair_mask(float c0, float c1, float c2, uint32_t bx, const ZCoefs& zc,
                                              const BlockCtx& b, uint32_t& cols) {
  const int l0 = bx & 31, l1 = (bx >> 5) & 31, l2 = (bx >> 10) & 31;
  const int h0 = (bx >> 15) & 31, h1 = (bx >> 20) & 31, h2 = (bx >> 25) & 31;
  const float lo0 = (float)(b.o0 + l0), hi0 = (float)(b.o0 + h0);
  float hmn[2], hmx[2], wmn[2], wmx[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    hmn[j] = wmn[j] = -INFINITY;
    hmx[j] = wmx[j] = INFINITY;
    if (zc.ok) {
      const int y0 = max(4 * j, l1), y1 = min(4 * j + 3, h1);
      const int x0 = max(4 * j, l2), x1 = min(4 * j + 3, h2);
      z_span(zc.h, (float)(b.o1 + y0) - c1, (float)(b.o1 + y1) - c1, hmn[j], hmx[j]);
      z_span(zc.w, (float)(b.o2 + x0) - c2, (float)(b.o2 + x1) - c2, wmn[j], wmx[j]);
    }
  }
  uint32_t mask = 0;
  cols = 0;
#pragma unroll
  for (int j1 = 0; j1 < 2; ++j1) {
#pragma unroll
    for (int j2 = 0; j2 < 2; ++j2) {
      if ((l1 >> 2) <= j1 && j1 <= (h1 >> 2) && (l2 >> 2) <= j2 && j2 <= (h2 >> 2)) {
        float zl, zh;
        column_rows(c0, hmn[j1], hmx[j1], wmn[j2], wmx[j2], lo0, hi0, zl, zh);
        if (zl <= zh) {
          const int za = (int)zl - b.o0, zb = (int)zh - b.o0;  // tile-local, inside [l0, h0]
          if (za <= 3) mask |= 1u << (2 * j1 + j2);
          if (zb >= 4) mask |= 1u << (4 + 2 * j1 + j2);
          cols |= col_field(za, zb) << (7 * (2 * j1 + j2));
        }
      }
    }
  }
  return mask;
}

// Appends the calling lanes' touching pairs to sh.touch (one LDS atomic per
// wave; every lane of the wave must call).
__device__ __forceinline__ void append_touch(Stage& sh, bool touch, int k, int lane) {
  const uint64_t m = __ballot(touch);
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int base = 0;
  if (lane == 0 && m != 0) base = atomicAdd(&sh.n_touch, (int)__popcll(m));
  base = __builtin_amdgcn_readfirstlane(base);
  if (touch) sh.touch[base + __popcll(m & lt_mask)] = (uint16_t)k;
}

// Sub-brick masks of the batch's touching pairs, spread over the block's
// threads (pairs that do not touch the tile cost nothing here).  Called
// between two barriers.  (k_forward, whose lists hold few pairs that miss the
// tile, computes each mask in the thread that staged the pair instead.)
__device__ __forceinline__ void batch_masks(Stage& sh, const BlockCtx& b, const ZRec* __restrict__ zt) {
  const int nt = sh.n_touch;
  for (int i = threadIdx.x; i < nt; i += kTileThreads) {
    const int k = sh.touch[i];
    uint32_t cols;
    sh.sbm[k] = (uint8_t)pair_mask(sh.c0[k], sh.c1[k], sh.c2[k], sh.box[k], z_load(zt, sh.gid[k]), b, cols);
  }
}

// Sub-brick mask of a pair staged at k from registers (its record r and
// column-bound record z; 0 unless it touches the tile).  Its column word (0
// unless it touches) replaces the box, which lists built from masks do not
// read; k_forward copies the words to pair_cols after the batch's evaluation.
__device__ __forceinline__ void stage_mask(Stage& sh, int k, const GRec& r, const ZRec& z, bool touch,
                                           const BlockCtx& b) {
  uint32_t cols;
  const uint32_t m = pair_mask(r.c0, r.c1, r.c2, sh.box[k], ZCoefs{z.h, z.w, z.h.rho > 0.0f}, b, cols);
  sh.sbm[k] = (uint8_t)(touch ? m : 0u);
  sh.box[k] = touch ? cols : 0u;
}

// Compacts, in batch order, the staged Gaussians whose box touches this
// wave's sub-brick into sh.list[wave]; returns the count (wave-uniform).
// Tiles of at most 2 sub-bricks per axis test one bit of the staged mask.
__device__ __forceinline__ int build_list(Stage& sh, int cnt, const BlockCtx& b, int wave, int lane, bool use_mask) {
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int nhit = 0;
  if (use_mask) {
    // The branch is uniform; kept outside the loop so that the loop holds
    // only the bit test.
    const int sbi = (b.s0 >> 2) * 4 + (b.s1 >> 2) * 2 + (b.s2 >> 2);
    for (int base = 0; base < cnt; base += 256) {
      // Lane l tests pairs base + 4l .. base + 4l + 3 from one 32-bit read;
      // their list slots follow those of the lower lanes (ascending pair order).
      const int k0 = base + 4 * lane;
      const uint32_t w4 = k0 < cnt ? *reinterpret_cast<const uint32_t*>(&sh.sbm[k0]) : 0u;
      bool h[4];
      int before = 0, total = 0;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        h[j] = k0 + j < cnt && ((w4 >> (8 * j + sbi)) & 1);
        const uint64_t m = __ballot(h[j]);
        before += __popcll(m & lt_mask);
        total += __popcll(m);
      }
      int pos = nhit + before;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (h[j]) sh.list[wave][pos] = (uint16_t)(k0 + j);
        pos += h[j] ? 1 : 0;
      }
      nhit += total;
    }
    return nhit;
  }
  for (int base = 0; base < cnt; base += 64) {
    const int k = base + lane;
    bool hit = false;
    if (k < cnt) {
      const uint32_t bx = sh.box[k];
      const int lo0 = bx & 31, lo1 = (bx >> 5) & 31, lo2 = (bx >> 10) & 31;
      const int hi0 = (bx >> 15) & 31, hi1 = (bx >> 20) & 31, hi2 = (bx >> 25) & 31;
      hit = !(bx & kEmptyBox) && lo0 < b.s0 + kSub && hi0 >= b.s0 && lo1 < b.s1 + kSub && hi1 >= b.s1 &&
            lo2 < b.s2 + kSub && hi2 >= b.s2;
    }
    const uint64_t m = __ballot(hit);
    if (hit) sh.list[wave][nhit + __popcll(m & lt_mask)] = (uint16_t)k;
    nhit += __popcll(m);
  }
  return nhit;
}

// Operations over the 16 lanes of each row of a wave with DPP lane moves (no
// LDS traffic).  All 64 lanes must be active.
template <int kCtrl>
__device__ __forceinline__ int dpp_mov(int v) {
  return __builtin_amdgcn_update_dpp(0, v, kCtrl, 0xf, 0xf, false);
}

// One halving step of a reduce-scatter over the 16 lanes of a row: lanes
// paired by the DPP pattern kCtrl (a mirror, so partners sit on opposite sides
// of the split), `upper` on the lane whose half is the upper one.  Each lane
// keeps the half of v it owns plus its partner's copy of that half.
template <int kCtrl, int kHalf>
__device__ __forceinline__ void row16_halve(float (&v)[2 * kHalf], bool upper) {
#pragma unroll
  for (int j = 0; j < kHalf; ++j) {
    const float send = upper ? v[j] : v[j + kHalf];
    const float keep = upper ? v[j + kHalf] : v[j];
    v[j] = keep + __int_as_float(dpp_mov<kCtrl>(__float_as_int(send)));
  }
}

// Reduce-scatter of 16 values over the 16 lanes of each row: lane l of the row
// ends with the sum over the row of value l in v[0].  Mirrors pair lane l
// with 15 - l (row_mirror), with 7 - l inside each half (row_half_mirror), with
// 3 - l and then l ^ 1 inside each quad; lane l's bits 3..0 pick the halves.
__device__ __forceinline__ void row16_reduce_scatter(float (&v)[16], int l) {
  row16_halve<0x140, 8>(v, (l & 8) != 0);
  float (&v8)[8] = reinterpret_cast<float (&)[8]>(v);
  row16_halve<0x141, 4>(v8, (l & 4) != 0);
  float (&v4)[4] = reinterpret_cast<float (&)[4]>(v);
  row16_halve<0x1B, 2>(v4, (l & 2) != 0);
  float (&v2)[2] = reinterpret_cast<float (&)[2]>(v);
  row16_halve<0xB1, 1>(v2, (l & 1) != 0);
}

// Maximum over the row, in every lane of the row: quad_perm [1,0,3,2] and
// [2,3,0,1] combine the lanes of each quad, then row_ror:4 and row_ror:8 the
// quads.
__device__ __forceinline__ int row16_max(int v) {
  v = max(v, dpp_mov<0xB1>(v));
  v = max(v, dpp_mov<0x4E>(v));
  v = max(v, dpp_mov<0x124>(v));
  v = max(v, dpp_mov<0x128>(v));
  return v;
}

// One lane's view of a Gaussian: the W offset u2 of its column is fixed, so
// the u2 terms of both linear forms and of the u2^2 term are per-lane values.
struct GaussLane {
  float c0, rr, ri, y1h, y1d, y2d;
  float g1, h2, c3;  // y1w u2, y2w u2, y3w u2^2
  float u1[kSub];    // H offsets of the 4 sub-brick rows
};

struct GaussRaw {
  float c0, c1, c2, rr, ri, y1h, y1d, y1w, y2d, y2w, y3w;
};

__device__ __forceinline__ GaussRaw load_raw(const Stage& sh, int k) {
  return GaussRaw{sh.c0[k], sh.c1[k], sh.c2[k], sh.rr[k], sh.ri[k], sh.y1h[k], sh.y1d[k], sh.y1w[k], sh.y2d[k],
                  sh.y2w[k], sh.y3w[k]};
}

__device__ __forceinline__ GaussLane lane_terms(const GaussRaw& r, const BlockCtx& b, int xi) {
  GaussLane g;
  g.c0 = r.c0;
  g.rr = r.rr;
  g.ri = r.ri;
  g.y1h = r.y1h;
  g.y1d = r.y1d;
  g.y2d = r.y2d;
  const float u2 = (float)(b.o2 + b.s2 + xi) - r.c2;
  g.g1 = r.y1w * u2;
  g.h2 = r.y2w * u2;
  g.c3 = r.y3w * u2 * u2;
#pragma unroll
  for (int y = 0; y < kSub; ++y) g.u1[y] = (float)(b.o1 + b.s1 + y) - r.c1;
  return g;
}

// Per D row u0: the H-independent part r = -(y2^2 + y3w u2^2) of q' and the
// part t1 of y1 without its H term; per voxel q' = -(y1h u1 + t1)^2 + r.
__device__ __forceinline__ void row_terms(const GaussLane& g, float u0, float& t1, float& r) {
  t1 = fmaf(g.y1d, u0, g.g1);
  const float y2 = fmaf(g.y2d, u0, g.h2);
  r = fmaf(-y2, y2, -g.c3);
}

// Evaluates the wave's compacted Gaussians 16 at a time over its sub-brick.
// kPartial: the sub-brick may extend past the tile's valid voxels (edge or
// non-multiple-of-4 tiles); full tiles take the branch-free path.
template <bool kPartial>
__device__ __forceinline__ void fwd_groups(const Stage& sh, int nhit, const BlockCtx& b, int wave, int gl, int xi,
                                           bool x_ok, float (&acc_r)[kSub * kSub], float (&acc_i)[kSub * kSub]) {
  for (int gb = 0; gb < nhit; gb += 16) {
    const int idx = gb + gl;
    const bool g_ok = (idx < nhit) && x_ok;
    const int k = sh.list[wave][idx < nhit ? idx : gb];
    const GaussLane g = lane_terms(load_raw(sh, k), b, xi);
#pragma unroll
    for (int z = 0; z < kSub; ++z) {
      if (!kPartial || z < b.n0) {
        const float u0 = (float)(b.o0 + b.s0 + z) - g.c0;
        float t1, r;
        row_terms(g, u0, t1, r);
#pragma unroll
        for (int y = 0; y < kSub; ++y) {
          if (!kPartial || y < b.n1) {
            const float y1 = fmaf(g.y1h, g.u1[y], t1);
            const float q = fmaf(-y1, y1, r);
            const bool ok = g_ok && (q >= kThr);
            const float w = ok ? __builtin_amdgcn_exp2f(q) : 0.0f;
            acc_r[z * kSub + y] = fmaf(w, g.rr, acc_r[z * kSub + y]);
            acc_i[z * kSub + y] = fmaf(w, g.ri, acc_i[z * kSub + y]);
          }
        }
      }
    }
  }
}

// Contributions of the tile's Gaussians whose density is not finite, after
// the batches (their staged density was zero, so the groups added nothing
// for them; a zero weight times inf would be NaN).  The tile's list is read
// again and such pairs staged in batches; thread (wave, gl, xi) evaluates the
// voxel (z, y, x) = (gl / 4, gl % 4, xi) of its wave's sub-brick -- the voxel
// whose total it holds in out_r / out_i -- and adds w rho only where the
// cutoff test passes, as the reference does.  The quadratic form is evaluated
// as in fwd_groups.  All threads of the block must call.
__device__ __forceinline__ void fwd_non_finite(Stage& sh, const GRec* __restrict__ rec,
                                               const int32_t* __restrict__ pair_ids, const BlockCtx& b, int gl,
                                               int xi, float& out_r, float& out_i) {
  const int z = gl >> 2, y = gl & 3;
  const bool vox_ok = b.active && z < b.n0 && y < b.n1 && xi < b.n2;
  const float x0 = (float)(b.o0 + b.s0 + z), x1 = (float)(b.o1 + b.s1 + y), x2 = (float)(b.o2 + b.s2 + xi);
  for (int bs = b.p_begin; bs < b.p_end; bs += kBatch) {
    const int cnt = min(kBatch, b.p_end - bs);
    __syncthreads();
    if (threadIdx.x == 0) sh.n_bad = 0;
    __syncthreads();
    if ((int)threadIdx.x < cnt) {
      const GRec r = rec[pair_ids[bs + threadIdx.x]];
      // e0 < 0 marks a non-finite centre (no voxel passes its test).
      if ((non_finite(r.rr) || non_finite(r.ri)) && r.e0 >= 0.0f) {
        const int j = atomicAdd(&sh.n_bad, 1);
        sh.c0[j] = r.c0;
        sh.c1[j] = r.c1;
        sh.c2[j] = r.c2;
        sh.rr[j] = r.rr;
        sh.ri[j] = r.ri;
        sh.y1h[j] = r.y1h;
        sh.y1d[j] = r.y1d;
        sh.y1w[j] = r.y1w;
        sh.y2d[j] = r.y2d;
        sh.y2w[j] = r.y2w;
        sh.y3w[j] = r.y3w;
      }
    }
    __syncthreads();
    const int nb = sh.n_bad;
    for (int j = 0; j < nb && vox_ok; ++j) {
      const float u0 = x0 - sh.c0[j], u1 = x1 - sh.c1[j], u2 = x2 - sh.c2[j];
      const float t1 = fmaf(sh.y1d[j], u0, sh.y1w[j] * u2);
      const float y2 = fmaf(sh.y2d[j], u0, sh.y2w[j] * u2);
      const float rz = fmaf(-y2, y2, -(sh.y3w[j] * u2 * u2));
      const float y1 = fmaf(sh.y1h[j], u1, t1);
      const float q = fmaf(-y1, y1, rz);
      if (q >= kThr) {
        const float w = __builtin_amdgcn_exp2f(q);
        out_r = fmaf(w, sh.rr[j], out_r);
        out_i = fmaf(w, sh.ri[j], out_i);
      }
    }
  }
}

// At most 128 VGPRs (4 waves per SIMD, two blocks per CU); the few values that
// do not fit are reloaded once per batch, outside the evaluation loops.
__global__ __launch_bounds__(kTileThreads, 4) void k_forward(const GRec* __restrict__ rec,
                                                             const ZRec* __restrict__ zt,
                                                             const int32_t* __restrict__ ranges,
                                                          const int32_t* __restrict__ pair_ids, VoxGeom geo,
                                                          float* __restrict__ vol_r, float* __restrict__ vol_i,
                                                          uint32_t* __restrict__ pair_cols) {
  __shared__ Stage sh;
  const int logical = xcd_remap(blockIdx.x, gridDim.x);
  if (logical >= geo.blocks) return;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int gl = lane & 15, xi = lane >> 4;
  const BlockCtx b = make_ctx(logical, geo, ranges, wave);
  if (!b.valid) return;
  const bool x_ok = xi < b.n2;
  const bool full = b.e0 == geo.sub && b.e1 == geo.sub && b.e2 == geo.sub && (geo.sub & 3) == 0;

  float acc_r[kSub * kSub], acc_i[kSub * kSub];
#pragma unroll
  for (int i = 0; i < kSub * kSub; ++i) {
    acc_r[i] = 0.0f;
    acc_i[i] = 0.0f;
  }

  // Records of the next batch are loaded into registers while the current
  // batch is evaluated (loads stay in flight across the LDS barriers).
  // Thread t stages pairs t and kBatch + t of a batch; of the second, only
  // the Gaussian id is read ahead.
  GRec nxt = {};
  int nxt_id = 0, nxt_id2 = 0;
  const bool use_mask = geo.nsb <= 2;
  if ((int)threadIdx.x < min(kBatch, b.p_end - b.p_begin)) {
    nxt_id = pair_ids[b.p_begin + threadIdx.x];
    nxt = load_rec<true>(rec, nxt_id);
  }
  if ((int)threadIdx.x + kBatch < min(kFwdBatch, b.p_end - b.p_begin)) nxt_id2 = pair_ids[b.p_begin + kBatch + threadIdx.x];
  if (threadIdx.x == 0) sh.any_bad = 0;
  for (int bs = b.p_begin; bs < b.p_end; bs += kFwdBatch) {
    const int cnt = min(kFwdBatch, b.p_end - bs);
    __syncthreads();
    // Thread t stages pair t (read during the previous batch) and pair
    // kBatch + t (read here), each with its column-bound record, and computes
    // their sub-brick masks.  A wave with any pair in a half stages that half
    // in all its lanes (lanes past the batch restage a record they hold into
    // slots no list reads), and a half's loads and staging share one
    // wave-uniform branch: every loaded value is then consumed on every path,
    // and the compiler keeps the next batch's loads in flight through the
    // evaluation instead of waiting for them in it.
    const int t2 = threadIdx.x + kBatch;
    auto stage_one = [&](int k, const GRec& r, const ZRec& z) {
      const bool touch = stage_pair(sh, k, r, r.gid, b) && k < cnt;
      if (use_mask) stage_mask(sh, k, r, z, touch, b);
      if (touch && (non_finite(r.rr) || non_finite(r.ri))) {
        // Zero weights times a non-finite density would be NaN: such a
        // Gaussian adds nothing here and is evaluated on its own after the
        // batches (fwd_non_finite).
        sh.rr[k] = 0.0f;
        sh.ri[k] = 0.0f;
        sh.any_bad = 1;
      }
    };
    if (kBatch + wave * 64 < cnt) {
      // Both halves: the second half's loads overlap the first's staging.
      ZRec z1 = {}, z2 = {};
      const GRec r2 = load_rec<true>(rec, nxt_id2);
      if (use_mask) {
        z2 = zt[nxt_id2];
        z1 = zt[nxt_id];
      }
      stage_one(threadIdx.x, nxt, z1);
      stage_one(t2, r2, z2);
    } else if (wave * 64 < cnt) {
      ZRec z1 = {};
      if (use_mask) z1 = zt[nxt_id];
      stage_one(threadIdx.x, nxt, z1);
    }
    const int bn = bs + kFwdBatch;
    if (bn < b.p_end) {
      if ((int)threadIdx.x < min(kBatch, b.p_end - bn)) {
        nxt_id = pair_ids[bn + threadIdx.x];
        nxt = load_rec<true>(rec, nxt_id);
      }
      if (t2 < min(kFwdBatch, b.p_end - bn)) nxt_id2 = pair_ids[bn + t2];
    }
    __syncthreads();
    if (b.active) {
      const int nhit = build_list(sh, cnt, b, wave, lane, use_mask);
      if (full)
        fwd_groups<false>(sh, nhit, b, wave, gl, xi, x_ok, acc_r, acc_i);
      else
        fwd_groups<true>(sh, nhit, b, wave, gl, xi, x_ok, acc_r, acc_i);
    }
    // Column words out after the evaluation: stores in flight during it would
    // make the compiler wait for them before reusing their data registers.
    if (use_mask && pair_cols != nullptr)
      for (int k = threadIdx.x; k < cnt; k += kTileThreads) pair_cols[bs + k] = sh.box[k];
  }

  // Sum the 16 Gaussian slots of each W column (a row of 16 lanes), scattered
  // so that lane gl of row xi holds position (z, y) = (gl / 4, gl % 4) of
  // column xi.
  row16_reduce_scatter(acc_r, gl);
  row16_reduce_scatter(acc_i, gl);
  float out_r = acc_r[0], out_i = acc_i[0];
  // any_bad is final: every batch's staging is behind a barrier.
  if (sh.any_bad != 0) fwd_non_finite(sh, rec, pair_ids, b, gl, xi, out_r, out_i);
  const int z = gl >> 2, y = gl & 3;
  if (b.active && x_ok && z < b.n0 && y < b.n1) {
    const int64_t flat = ((int64_t)(b.o0 + b.s0 + z) * geo.H + (b.o1 + b.s1 + y)) * geo.W + (b.o2 + b.s2 + xi);
    vol_r[flat] = out_r;
    vol_i[flat] = out_i;
  }
}

// Cross-row sums with the gfx950 lane-swap instructions.  A wave is four
// rows of 16 lanes.  v_permlane32_swap(a, b) exchanges the upper half of a with
// the lower half of b, so a' + b' holds, in the lower half, a summed over rows
// {r, r+2} and, in the upper half, b summed the same way.
__device__ __forceinline__ float pair_sum_rows02(float a, float b) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

// v_permlane16_swap(a, b) exchanges the odd rows of a with the even rows of b,
// so a' + b' holds a summed over rows {0,1} in row 0, b over {0,1} in row 1,
// a over {2,3} in row 2 and b over {2,3} in row 3.
__device__ __forceinline__ float pair_sum_rows01(float a, float b) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

constexpr int kMoments = 11;  // rho_r, rho_i, 3 first moments, 6 second moments
constexpr int kAccStride = 16;  // floats per accumulator row: one 64-byte row per Gaussian

// ---- backward ----
//
// The backward gives every Gaussian of a batch to one set of lanes for the
// whole tile, so its 11 moments accumulate in registers and no LDS atomics are
// needed inside the tile.  A lane = (Gaussian slot, W offset) walks the live
// columns of its Gaussian -- the 4 x 4 (H, W) voxel blocks of its box that the
// column bound above leaves non-empty -- and in each column exactly the D rows
// the bound leaves.  A Gaussian's columns are kept in (H, W) block order (the
// order in which the 16 lanes' upstream-gradient reads spread best over the
// LDS banks); Gaussians are keyed by (live columns, largest row count) and
// counting-sorted, a group is 16 Gaussians with the same column count and
// neighbouring row counts, and the group runs its j-th columns for the largest
// row count among them.  The 8 waves take groups from a block-wide counter in
// descending cost order, so a wave that finishes early takes the next group.

constexpr int kMaxBlk = 4;                  // column blocks per axis, tile <= 16
constexpr int kMaxCols = kMaxBlk * kMaxBlk;  // columns per tile
constexpr int kMaxRows = 16;                // D rows for tile <= 16
constexpr int kKeys = kMaxCols * kMaxRows;
constexpr uint16_t kNoKey = 0xffff;

// Column word: H block (bits 0-1), W block (2-3), first row (4-7), rows - 1 (8-11).
__device__ __forceinline__ uint16_t pack_col(int j1, int j2, int z0, int rows) {
  return (uint16_t)(j1 | (j2 << 2) | (z0 << 4) | ((rows - 1) << 8));
}
__device__ __forceinline__ int col_rows(uint16_t c) { return ((c >> 8) & 15) + 1; }

// Strides (float2 elements) of the tile's upstream-gradient copy in LDS: rows
// padded by 2 and slabs by 2 elements spread the 16 Gaussians' column reads of
// a wave over the banks (a sweep of row/slab paddings on the microbenchmark
// found this among the fastest, ~1 % ahead of slabs padded by 4).
__host__ __device__ constexpr int grad_ys(int t) { return t + 2; }
__host__ __device__ constexpr int grad_zs(int t) { return t * (t + 2) + 2; }

// MC: live columns a pair can have (column blocks per tile).
// Pairs staged per batch of the backward.  8-voxel tiles stage up to two per
// thread: most tiles of the microbenchmark's fused-call lists (≈ 580 pairs per
// tile) then take one batch and their Gaussians are grouped together.  Larger
// tiles keep 384, so that their bigger LDS gradient copy still leaves room for
// two blocks per CU.
template <int TT>
constexpr int bwd_batch() {
  return TT == 8 ? 2 * kTileThreads : 384;
}

template <int MC, int NB>
struct BwdStage {
  float c0[NB], c1[NB], c2[NB], rr[NB], ri[NB];
  float y1h[NB], y1d[NB], y1w[NB], y2d[NB], y2w[NB], y3w[NB];
  int32_t gid[NB];            // Gaussian id of each staged pair
  uint16_t cols[NB][MC];      // live columns in (H, W) block order
  uint16_t key[NB];           // (columns - 1) * 16 + largest row count - 1, or kNoKey
  uint16_t order[NB];         // batch indices sorted by key
  uint16_t slot[NB];          // rank inside the key
  int key_cnt[kKeys];
  int key_start[kKeys];
  int seg_begin[kMaxCols];        // per column count cc (index cc - 1): first sorted position
  int seg_cnt[kMaxCols];          //   number of Gaussians
  int grp_first[kMaxCols];        //   first group; groups run in descending column count
  int n_groups;                   // groups of the batch
  int next_group;                 // next group to take
};

struct BwdAcc {
  float dr, di, m, m0, m00, m1, m01, m11, m2, m22, m02, m12;
};

// One column (4 x 4 block of H, W) of one lane's Gaussian: `rows` D rows from
// tile-local z0 (rows beyond the lane's own `own_rows` contribute nothing).
// TT is the tile edge when known at compile time (0: runtime).
// LDS-qualified view of the tile's upstream-gradient copy (32-bit addressing),
// as interleaved (real, imag) floats.
typedef __attribute__((address_space(3))) const float LdsFloat;

__device__ __forceinline__ float2 lds_g(LdsFloat* __restrict__ g, int i) { return make_float2(g[2 * i], g[2 * i + 1]); }

// kSafe: the group's densities or the tile's upstream gradient hold a
// non-finite value; contributions are then selected by the cutoff test
// instead of weighted by a zero (0 * inf is NaN), so, as in the reference,
// only evaluations with q <= 9 add anything.
template <int TT, bool kPartial, bool kSafe>
__device__ __forceinline__ void bwd_column(LdsFloat* __restrict__ sh_g, int tp, const BlockCtx& b, int z0,
                                           int rows, int own_rows, int s1, int s2, int xi, bool g_ok, float c0,
                                           float c1, float c2, float rr, float ri, float y1h, float y1d, float y1w,
                                           float y2d, float y2w, float y3w, BwdAcc& acc) {
  const int T = TT > 0 ? TT : tp;
  const int YS = grad_ys(T), ZS = grad_zs(T);
  const float u2 = (float)(b.o2 + s2 + xi) - c2;
  const float g1 = y1w * u2, h2 = y2w * u2, c3 = y3w * u2 * u2;
  const int gbase = s1 * YS + s2 + xi;
  bool ok_lane = g_ok;
  if (kPartial) ok_lane = ok_lane && (s2 + xi < b.e2);
  float m = 0.0f, m0 = 0.0f, m1 = 0.0f;
  float u1v[kSub], u1q[kSub];
#pragma unroll
  for (int y = 0; y < kSub; ++y) {
    u1v[y] = (float)(b.o1 + s1 + y) - c1;
    u1q[y] = u1v[y] * u1v[y];
  }
  // The next row's upstream gradient is loaded from LDS while the current row
  // is evaluated.  With the zero slabs past the tile (kZeroSlabs), rows past
  // the lane's own need no mask: inside the tile the column bound leaves no
  // voxel there whose q passes the cutoff, and past it the gradient is zero
  // (the weights are at most 1).  Otherwise, and on the select path, where a
  // density may be infinite, they get q = -inf and re-read the lane's last row.
  constexpr bool kZeroSlabs = TT == 8;
  const int zlast_off = (z0 + own_rows - 1) * ZS;
  int zn_off = (z0 + 1) * ZS;
  float2 gn[kSub];
#pragma unroll
  for (int y = 0; y < kSub; ++y) gn[y] = lds_g(sh_g, gbase + z0 * ZS + y * YS);
  float u0 = (float)(b.o0 + z0) - c0;
#pragma unroll 2
  for (int r = 0; r < rows; ++r) {
    // q' = -(y1h u1 + p1)^2 + rz: rz = -(y2^2 + y3w u2^2), and -inf on rows
    // past the lane's own (no voxel passes).
    const float p1 = fmaf(y1d, u0, g1);
    const float y2 = fmaf(y2d, u0, h2);
    float rz = fmaf(-y2, y2, -c3);
    if (!kZeroSlabs || kSafe) rz = r < own_rows ? rz : -INFINITY;
    float2 gc[kSub];
#pragma unroll
    for (int y = 0; y < kSub; ++y) gc[y] = gn[y];
    {
      const int nb = gbase + ((kZeroSlabs && !kSafe) ? zn_off : min(zn_off, zlast_off));
      zn_off += ZS;
#pragma unroll
      for (int y = 0; y < kSub; ++y) gn[y] = lds_g(sh_g, nb + y * YS);
    }
    float wt[kSub];
#pragma unroll
    for (int y = 0; y < kSub; ++y) {
      float ry = rz;
      if (kPartial) ry = (s1 + y < b.e1) ? rz : -INFINITY;
      const float y1 = fmaf(y1h, u1v[y], p1);
      const float q = fmaf(-y1, y1, ry);
      const bool ok = ok_lane && (q >= kThr);
      const float w = ok ? __builtin_amdgcn_exp2f(q) : 0.0f;
      const float2 g = gc[y];
      if (kSafe) {
        acc.dr = ok ? fmaf(w, g.x, acc.dr) : acc.dr;
        acc.di = ok ? fmaf(w, g.y, acc.di) : acc.di;
        wt[y] = ok ? w * fmaf(ri, g.y, rr * g.x) : 0.0f;
      } else {
        acc.dr = fmaf(w, g.x, acc.dr);
        acc.di = fmaf(w, g.y, acc.di);
        wt[y] = w * fmaf(ri, g.y, rr * g.x);
      }
    }
    // Row moments t0 = sum wt, t1 = sum wt u1, t2 = sum wt u1^2 from each
    // voxel's own offset (exact in fp32) and its square: forming them from
    // the first row's offset, u1 = u10 + y, would leave an absolute error of
    // ~ulp(u10) sum|wt| that cov_inv magnifies in the centre gradient of
    // Gaussians much narrower than a voxel.
    const float t0 = (wt[0] + wt[1]) + (wt[2] + wt[3]);
    const float t1 = fmaf(wt[3], u1v[3], fmaf(wt[2], u1v[2], fmaf(wt[1], u1v[1], wt[0] * u1v[0])));
    const float t2 = fmaf(wt[3], u1q[3], fmaf(wt[2], u1q[2], fmaf(wt[1], u1q[1], wt[0] * u1q[0])));
    m += t0;
    const float tz = t0 * u0;
    m0 += tz;
    acc.m00 = fmaf(tz, u0, acc.m00);
    m1 += t1;
    acc.m01 = fmaf(t1, u0, acc.m01);
    acc.m11 += t2;
    u0 += 1.0f;
  }
  acc.m += m;
  acc.m0 += m0;
  acc.m1 += m1;
  acc.m2 = fmaf(u2, m, acc.m2);
  acc.m22 = fmaf(u2 * u2, m, acc.m22);
  acc.m02 = fmaf(u2, m0, acc.m02);
  acc.m12 = fmaf(u2, m1, acc.m12);
}

// pair_cols (optional, tiles of at most 8 voxels per axis): the forward's
// column words of the same pairs, used instead of recomputing the bound.
template <int TT>
__global__ __launch_bounds__(kTileThreads, 4) void k_backward(const GRec* __restrict__ rec,
                                                           const ZRec* __restrict__ zt,
                                                           const int32_t* __restrict__ ranges,
                                                           const int32_t* __restrict__ pair_ids, VoxGeom geo,
                                                           const float* __restrict__ grad_r,
                                                           const float* __restrict__ grad_i,
                                                           float* __restrict__ gacc,
                                                           const uint32_t* __restrict__ pair_cols) {
  constexpr int NB = TT > 0 ? (TT + kSub - 1) / kSub : kMaxBlk;  // column blocks per axis (bound)
  constexpr int MC = NB * NB;
  constexpr int kB = bwd_batch<TT>();
  __shared__ BwdStage<MC, kB> sh;
  // Upstream gradient of the tile, TP^3 with TP the tile edge rounded up to a
  // multiple of 4 (strides grad_zs, grad_ys, 1), zero outside the tile's valid
  // voxels.
  extern __shared__ float2 sh_g[];
  const int logical = xcd_remap(blockIdx.x, gridDim.x);
  if (logical >= geo.blocks) return;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int gl = lane & 15, xi = lane >> 4;
  const int T = TT > 0 ? TT : geo.nsb * kSub;
  const int nb = TT > 0 ? NB : geo.nsb;  // column blocks per axis
  // One block per tile: the context's wave-specific fields are unused here.
  const BlockCtx b = make_ctx(logical, geo, ranges, 0);
  if (!b.valid) return;
  const bool partial = b.e1 < T || b.e2 < T;  // D rows are always clipped to the tile

  // Records of the next batch are loaded into registers while the current
  // batch is evaluated (loads stay in flight across the LDS barriers); the
  // first batch's are requested before the gradient copy, so that the two
  // reads overlap.
  const bool use_cols = pair_cols != nullptr && nb <= 2;
  const int bsz = kB;
  // Thread t stages pairs t and kTileThreads + t of a batch; of the second,
  // only the Gaussian id is read ahead.
  GRec nxt = {};
  int nxt_id = 0, nxt_id2 = 0;
  uint32_t nxt_cols = 0;
  if (tid < min(bsz, b.p_end - b.p_begin)) {
    nxt_id = pair_ids[b.p_begin + tid];
    nxt = load_rec<false>(rec, nxt_id);
    if (use_cols) nxt_cols = pair_cols[b.p_begin + tid];
  }
  if (tid + kTileThreads < min(bsz, b.p_end - b.p_begin)) nxt_id2 = pair_ids[b.p_begin + kTileThreads + tid];

  int g_bad = 0;
  // For 8-voxel tiles the copy has 8 more slabs of zeros past the tile: a
  // group's row walk from any first row stays inside it.
  const int copy_slabs = TT == 8 ? 2 * T : T;
  for (int i = tid; i < copy_slabs * T * T; i += kTileThreads) {
    const int z = i / (T * T), y = (i / T) % T, x = i % T;
    float2 g = make_float2(0.0f, 0.0f);
    if (z < b.e0 && y < b.e1 && x < b.e2) {
      const int64_t flat = ((int64_t)(b.o0 + z) * geo.H + (b.o1 + y)) * geo.W + (b.o2 + x);
      g = make_float2(grad_r[flat], grad_i[flat]);
      g_bad |= (non_finite(g.x) || non_finite(g.y)) ? 1 : 0;
    }
    sh_g[z * grad_zs(T) + y * grad_ys(T) + x] = g;
  }
  for (int i = tid; i < kKeys; i += kTileThreads) sh.key_cnt[i] = 0;
  // A non-finite upstream gradient anywhere in the tile sends every group of
  // the tile down the select path (block-uniform).
  const bool tile_bad = __syncthreads_or(g_bad) != 0;

  for (int bs = b.p_begin; bs < b.p_end; bs += bsz) {
    const int cnt = min(bsz, b.p_end - bs);
    __syncthreads();
    // Stage: parameters, live columns and key of every pair of the batch;
    // thread t stages pair t (loaded during the previous batch) and pair
    // kTileThreads + t (read here).
    if (tid == 0) sh.next_group = 0;
    auto stage = [&](int t, const GRec& r, int id, uint32_t colw) {
      sh.gid[t] = id;
      sh.c0[t] = r.c0;
      sh.c1[t] = r.c1;
      sh.c2[t] = r.c2;
      sh.rr[t] = r.rr;
      sh.ri[t] = r.ri;
      sh.y1h[t] = r.y1h;
      sh.y1d[t] = r.y1d;
      sh.y1w[t] = r.y1w;
      sh.y2d[t] = r.y2d;
      sh.y2w[t] = r.y2w;
      sh.y3w[t] = r.y3w;
      uint16_t key = kNoKey;
      const float lo0 = fmaxf(ceilf(r.c0 - r.e0), (float)b.o0);
      const float lo1 = fmaxf(ceilf(r.c1 - r.e1), (float)b.o1);
      const float lo2 = fmaxf(ceilf(r.c2 - r.e2), (float)b.o2);
      const float hi0 = fminf(floorf(r.c0 + r.e0), (float)(b.o0 + b.e0 - 1));
      const float hi1 = fminf(floorf(r.c1 + r.e1), (float)(b.o1 + b.e1 - 1));
      const float hi2 = fminf(floorf(r.c2 + r.e2), (float)(b.o2 + b.e2 - 1));
      const bool touch = lo0 <= hi0 && lo1 <= hi1 && lo2 <= hi2;
      if (use_cols && touch) {
        // Live columns from the forward's column word (written for the pairs
        // whose box touches the tile), in (H, W) block order.
        uint16_t* cl = sh.cols[t];
        int ncol = 0, mxr = 0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const uint32_t f = (colw >> (7 * c)) & 127u;
          if (f & 1u) {
            const int rows = (int)((f >> 4) & 7u) + 1;
            cl[ncol] = pack_col(c >> 1, c & 1, (int)((f >> 1) & 7u), rows);
            ++ncol;
            mxr = max(mxr, rows);
          }
        }
        if (ncol > 0) {
          key = (uint16_t)((ncol - 1) * kMaxRows + mxr - 1);
          sh.slot[t] = (uint16_t)atomicAdd(&sh.key_cnt[key], 1);
        }
      }
      if (!use_cols && touch) {
        const int y_lo = (int)lo1 - b.o1, y_hi = (int)hi1 - b.o1;  // tile-local box
        const int x_lo = (int)lo2 - b.o2, x_hi = (int)hi2 - b.o2;
        const int l1 = y_lo >> 2, h1 = y_hi >> 2, l2 = x_lo >> 2, h2 = x_hi >> 2;
        const ZCoefs zc = z_load(zt, id);
        float hmn[NB], hmx[NB], wmn[NB], wmx[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j) {
          hmn[j] = wmn[j] = -INFINITY;
          hmx[j] = wmx[j] = INFINITY;
          if (zc.ok) {
            const int y0 = max(4 * j, y_lo), y1 = min(4 * j + 3, y_hi);
            const int x0 = max(4 * j, x_lo), x1 = min(4 * j + 3, x_hi);
            z_span(zc.h, (float)(b.o1 + y0) - r.c1, (float)(b.o1 + y1) - r.c1, hmn[j], hmx[j]);
            z_span(zc.w, (float)(b.o2 + x0) - r.c2, (float)(b.o2 + x1) - r.c2, wmn[j], wmx[j]);
          }
        }
        // Live columns in (H, W) block order, and their largest row count.
        uint16_t* cl = sh.cols[t];
        int ncol = 0;
        int mxr = 0;
#pragma unroll
        for (int j1 = 0; j1 < NB; ++j1) {
#pragma unroll
          for (int j2 = 0; j2 < NB; ++j2) {
            if (j1 < nb && j2 < nb && l1 <= j1 && j1 <= h1 && l2 <= j2 && j2 <= h2) {
              float zl, zh;
              column_rows(r.c0, hmn[j1], hmx[j1], wmn[j2], wmx[j2], lo0, hi0, zl, zh);
              if (zl <= zh) {
                const int rows = (int)zh - (int)zl + 1;
                const uint16_t c = pack_col(j1, j2, (int)zl - b.o0, rows);
                cl[ncol] = c;
                ++ncol;
                mxr = max(mxr, rows);
              }
            }
          }
        }
        if (ncol > 0) {
          key = (uint16_t)((ncol - 1) * kMaxRows + mxr - 1);
          sh.slot[t] = (uint16_t)atomicAdd(&sh.key_cnt[key], 1);
        }
      }
      sh.key[t] = key;
    };
    if (tid < cnt) stage(tid, nxt, nxt_id, nxt_cols);
    if (tid + kTileThreads < cnt) {
      const int t2 = tid + kTileThreads;
      stage(t2, load_rec<false>(rec, nxt_id2), nxt_id2, use_cols ? pair_cols[bs + t2] : 0u);
    }
    {
      const int bn = bs + bsz;
      if (bn < b.p_end && tid < min(bsz, b.p_end - bn)) {
        nxt_id = pair_ids[bn + tid];
        nxt = load_rec<false>(rec, nxt_id);
        if (use_cols) nxt_cols = pair_cols[bn + tid];
      }
      if (bn < b.p_end && tid + kTileThreads < min(bsz, b.p_end - bn)) nxt_id2 = pair_ids[bn + kTileThreads + tid];
    }
    __syncthreads();
    // Key offsets (ascending keys) and, per column count, the sorted segment and
    // its first group (groups run in descending column count).
    if (wave == 0) {
      int c[4], sum = 0;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        c[e] = sh.key_cnt[lane * 4 + e];
        sum += c[e];
      }
      int inc = sum;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int o = __shfl_up(inc, off, 64);
        if (lane >= off) inc += o;
      }
      int run = inc - sum;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        sh.key_start[lane * 4 + e] = run;
        run += c[e];
      }
      // Column count cc owns keys [16 (cc - 1), 16 cc), i.e. lanes 4 (cc - 1) .. 4 (cc - 1) + 3.
      const int seg = sum + __shfl_down(sum, 1, 64) + __shfl_down(sum, 2, 64) + __shfl_down(sum, 3, 64);
      const int ng = (lane & 3) == 0 ? (seg + 15) >> 4 : 0;
      int gin = ng;  // suffix sum over the segment lanes (stride 4)
#pragma unroll
      for (int off = 4; off < 64; off <<= 1) {
        const int o = __shfl_down(gin, off, 64);
        if (lane + off < 64) gin += o;
      }
      if ((lane & 3) == 0) {
        sh.seg_begin[lane >> 2] = inc - sum;
        sh.seg_cnt[lane >> 2] = seg;
        sh.grp_first[lane >> 2] = gin - ng;
      }
      if (lane == 0) sh.n_groups = gin;
    }
    __syncthreads();
    for (int t = tid; t < cnt; t += kTileThreads)
      if (sh.key[t] != kNoKey) sh.order[sh.key_start[sh.key[t]] + sh.slot[t]] = (uint16_t)t;
    __syncthreads();

    // Groups are taken in descending cost order (column count, then rows) from
    // a block-wide counter, so a wave that finishes early takes the next one.
    const int n_groups = sh.n_groups;
#pragma unroll 1
    for (;;) {
      int g_next = 0;
      if (lane == 0) g_next = atomicAdd(&sh.next_group, 1);
      g_next = __builtin_amdgcn_readfirstlane(g_next);
      if (g_next >= n_groups) break;
      // Its segment (column count cc): grp_first[cc-1] <= g < grp_first[cc-1] + groups(cc).
      int cc = MC;
      while (g_next >= sh.grp_first[cc - 1] + ((sh.seg_cnt[cc - 1] + 15) >> 4)) --cc;
      const int seg_begin = sh.seg_begin[cc - 1];
      const int seg_cnt = sh.seg_cnt[cc - 1];
      {
        const int r = g_next - sh.grp_first[cc - 1];
        // Group r holds sorted positions top - 15 .. top (largest keys first).
        const int top = seg_begin + seg_cnt - 1 - 16 * r;
        const int pos = top - gl;
        const bool slot_ok = pos >= seg_begin;
        const int k = sh.order[slot_ok ? pos : top];
        const float c0 = sh.c0[k], c1 = sh.c1[k], c2 = sh.c2[k], rr = sh.rr[k], ri = sh.ri[k];
        const float y1h = sh.y1h[k], y1d = sh.y1d[k], y1w = sh.y1w[k];
        const float y2d = sh.y2d[k], y2w = sh.y2w[k], y3w = sh.y3w[k];
        BwdAcc acc = {};
        // Lanes of empty slots repeat the group's top Gaussian, so the
        // wave-wide test covers every density the group multiplies.
        const bool safe = tile_bad || __ballot(non_finite(rr) || non_finite(ri)) != 0;
        for (int j = 0; j < cc; ++j) {
          const uint16_t cw = sh.cols[k][j];
          const int own_rows = col_rows(cw);
          // The group's largest row count for its j-th columns (over the 16 slots;
          // every row of the wave holds the same 16 Gaussians, so it is uniform
          // and the row loop counts in scalar registers).
          const int rows = __builtin_amdgcn_readfirstlane(row16_max(own_rows));
          const int j1 = cw & 3, j2 = (cw >> 2) & 3, z0 = (cw >> 4) & 15;
          LdsFloat* lg = (LdsFloat*)sh_g;
          if (safe) {
            if (partial)
              bwd_column<TT, true, true>(lg, T, b, z0, rows, own_rows, j1 * kSub, j2 * kSub, xi, slot_ok, c0, c1, c2,
                                         rr, ri, y1h, y1d, y1w, y2d, y2w, y3w, acc);
            else
              bwd_column<TT, false, true>(lg, T, b, z0, rows, own_rows, j1 * kSub, j2 * kSub, xi, slot_ok, c0, c1,
                                          c2, rr, ri, y1h, y1d, y1w, y2d, y2w, y3w, acc);
          } else if (partial) {
            bwd_column<TT, true, false>(lg, T, b, z0, rows, own_rows, j1 * kSub, j2 * kSub, xi, slot_ok, c0, c1, c2,
                                        rr, ri, y1h, y1d, y1w, y2d, y2w, y3w, acc);
          } else {
            bwd_column<TT, false, false>(lg, T, b, z0, rows, own_rows, j1 * kSub, j2 * kSub, xi, slot_ok, c0, c1, c2,
                                         rr, ri, y1h, y1d, y1w, y2d, y2w, y3w, acc);
          }
        }
        const float v[12] = {acc.dr,  acc.di,  acc.m0,  acc.m1,  acc.m2,  acc.m00,
                             acc.m11, acc.m22, acc.m01, acc.m02, acc.m12, 0.0f};
        // Reduce-scatter over the 4 W lanes of each Gaussian: afterwards the
        // lane in row xi holds the totals of moments 4*jj + xi, jj = 0..2.
        float t3[3];
#pragma unroll
        for (int jj = 0; jj < 3; ++jj) {
          const float lo = pair_sum_rows02(v[4 * jj], v[4 * jj + 2]);      // rows 0,1: m 4jj; rows 2,3: m 4jj+2
          const float hi = pair_sum_rows02(v[4 * jj + 1], v[4 * jj + 3]);  // rows 0,1: m 4jj+1; rows 2,3: m 4jj+3
          t3[jj] = pair_sum_rows01(lo, hi);
        }
        // The totals go to the Gaussians' 64-byte rows by whole rows: atomic
        // instruction i covers Gaussians 4i .. 4i + 3 of the group, lane L
        // adding moment L % 16 of Gaussian 4i + L / 16, so that it touches 4
        // cache lines instead of 16 (the L2 handles a float atomic instruction
        // per line it touches).  Moment m of Gaussian g sits in register t3[m / 4]
        // of lane 16 (m % 4) + g; lane g holds the Gaussian's id.  The atomics
        // drain while the wave evaluates its next group.
        {
          const int m = lane & 15, jm = m >> 2;
          const int id = slot_ok ? sh.gid[k] : -1;
          // All lane moves first, then the atomics: one wait for the moves.
          float val[4];
          int gid[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int g = 4 * i + (lane >> 4);
            const int src = ((m & 3) * 16 + g) * 4;
            const float a0 = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(t3[0])));
            const float a1 = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(t3[1])));
            const float a2 = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(t3[2])));
            val[i] = jm == 0 ? a0 : (jm == 1 ? a1 : a2);
            gid[i] = __builtin_amdgcn_ds_bpermute(g * 4, id);
          }
#pragma unroll
          for (int i = 0; i < 4; ++i)
            if (m < kMoments && gid[i] >= 0 && val[i] != 0.0f)
              unsafeAtomicAdd(gacc + (int64_t)gid[i] * kAccStride + m, val[i]);
        }
      }
    }
    // key_cnt is only read before the group loop; the next batch's barrier
    // orders this reset before its counts.
    for (int i = tid; i < kKeys; i += kTileThreads) sh.key_cnt[i] = 0;
  }
}

// Moments -> reference outputs (gaussian/cuda_ext/tile_voxelize_cuda_kernel.cu:237-252):
//   s = -w t / 2;  grad_rho = sum w g;  grad_c_i = -2/hs_i sum_j M_ij sum(s u_j)/hs_j;
//   grad_M_ij = sum(s u_i u_j) / (hs_i hs_j).
// One Gaussian's moments -> grad_center (gc), grad_cov_inv (gm), grad_rho.
// In fp64: for strongly anisotropic Gaussians M (m / hs) cancels -- M's large
// eigenvalue multiplies the small component of m along its eigenvector -- and
// an fp32 product would add its own rounding to that of the fp32 moments.
__device__ __forceinline__ void finalize_one(const float* __restrict__ gacc, const float* __restrict__ cov,
                                             const float hs[3], int g, double gc[3], double gm[9], float& grr,
                                             float& gri) {
  const float* a = gacc + (int64_t)g * kAccStride;
  grr = a[0];
  gri = a[1];
  const double m[3] = {-0.5 * a[2], -0.5 * a[3], -0.5 * a[4]};
  const double mm[3][3] = {{-0.5 * a[5], -0.5 * a[8], -0.5 * a[9]},
                           {-0.5 * a[8], -0.5 * a[6], -0.5 * a[10]},
                           {-0.5 * a[9], -0.5 * a[10], -0.5 * a[7]}};
  const double h[3] = {hs[0], hs[1], hs[2]};
  const float* M = cov + 9 * (int64_t)g;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) acc = fma((double)M[3 * i + j], m[j] / h[j], acc);
    gc[i] = -2.0 * acc / h[i];
#pragma unroll
    for (int j = 0; j < 3; ++j) gm[3 * i + j] = mm[i][j] / (h[i] * h[j]);
  }
  // The reference accumulates the centre gradient over the voxels that pass
  // the cutoff only: with none (zero first moments) it is exactly zero, also
  // where cov_inv is not finite (then no voxel passes).
  if (a[2] == 0.0f && a[3] == 0.0f && a[4] == 0.0f) gc[0] = gc[1] = gc[2] = 0.0;
}

__global__ __launch_bounds__(256) void k_finalize(const float* __restrict__ gacc, const float* __restrict__ cov,
                                                  const float* __restrict__ half_shape, int n,
                                                  float* __restrict__ g_center, float* __restrict__ g_cov,
                                                  float* __restrict__ g_rr, float* __restrict__ g_ri,
                                                  int rho_stride) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= n) return;
  const float hs[3] = {half_shape[0], half_shape[1], half_shape[2]};
  double gc[3], gm[9];
  float grr, gri;
  finalize_one(gacc, cov, hs, g, gc, gm, grr, gri);
  g_rr[(int64_t)g * rho_stride] = grr;
  g_ri[(int64_t)g * rho_stride] = gri;
#pragma unroll
  for (int i = 0; i < 3; ++i) g_center[3 * (int64_t)g + i] = (float)gc[i];
#pragma unroll
  for (int i = 0; i < 9; ++i) g_cov[9 * (int64_t)g + i] = (float)gm[i];
}

// ------------------------------------------------------ prologue / epilogue --

// NaN-propagating maximum (torch.max semantics).
__device__ __forceinline__ float max_nan(float a, float b) { return (a != a || a >= b) ? a : b; }

// Fused prologue of TileVoxelizer.forward (gaussian/tile_voxelizer.py:281-290):
//   hs = shape/2;  c = ((p + 1) * 0.5) * shape - 0.5;  r = max_a(s_a * hs_a) * 3;
//   R = R(q / (|q| + 1e-8));  T = R * (1 / (s + 1e-8)) (column scaling);  M = T T^T.
// Centre and radius round after every operation (no contraction) so binning
// sees the same fp32 values the reference's torch ops produce.
// Prologue of Gaussian g (gaussian/tile_voxelizer.py:281-290): writes its
// centre, radius and cov_inv and returns the centre and matrix in c and M.
// Every operation rounds (no contraction): centre and radius match the
// reference's fp32 operation order bit for bit.
__device__ __forceinline__ void prologue_one(int g, const float* __restrict__ pos, const float* __restrict__ scales,
                                             const float* __restrict__ rot, float sh0, float sh1, float sh2,
                                             float* __restrict__ center, float* __restrict__ radius,
                                             float* __restrict__ cov, float c[3], float M[9], float& rad) {
#pragma clang fp contract(off)
  const float hs0 = sh0 * 0.5f, hs1 = sh1 * 0.5f, hs2 = sh2 * 0.5f;
  const int64_t g3 = 3 * (int64_t)g;
  c[0] = ((pos[g3 + 0] + 1.0f) * 0.5f) * sh0 - 0.5f;
  c[1] = ((pos[g3 + 1] + 1.0f) * 0.5f) * sh1 - 0.5f;
  c[2] = ((pos[g3 + 2] + 1.0f) * 0.5f) * sh2 - 0.5f;
  center[g3 + 0] = c[0];
  center[g3 + 1] = c[1];
  center[g3 + 2] = c[2];
  const float s0 = scales[g3 + 0], s1 = scales[g3 + 1], s2 = scales[g3 + 2];
  rad = max_nan(max_nan(s0 * hs0, s1 * hs1), s2 * hs2) * 3.0f;
  radius[g] = rad;
  const float4 q4 = *reinterpret_cast<const float4*>(rot + 4 * (int64_t)g);
  const float nrm = sqrtf(q4.x * q4.x + q4.y * q4.y + q4.z * q4.z + q4.w * q4.w) + 1e-8f;
  const float w = q4.x / nrm, x = q4.y / nrm, y = q4.z / nrm, z = q4.w / nrm;
  const float R[3][3] = {{1.0f - 2.0f * y * y - 2.0f * z * z, 2.0f * x * y - 2.0f * z * w, 2.0f * x * z + 2.0f * y * w},
                         {2.0f * x * y + 2.0f * z * w, 1.0f - 2.0f * x * x - 2.0f * z * z, 2.0f * y * z - 2.0f * x * w},
                         {2.0f * x * z - 2.0f * y * w, 2.0f * y * z + 2.0f * x * w, 1.0f - 2.0f * x * x - 2.0f * y * y}};
  const float si[3] = {1.0f / (s0 + 1e-8f), 1.0f / (s1 + 1e-8f), 1.0f / (s2 + 1e-8f)};
  float T[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) T[i][k] = R[i][k] * si[k];
  float* Mg = cov + 9 * (int64_t)g;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      M[3 * i + j] = T[i][0] * T[j][0] + T[i][1] * T[j][1] + T[i][2] * T[j][2];
      Mg[3 * i + j] = M[3 * i + j];
    }
}

__global__ __launch_bounds__(256) void k_prologue(const float* __restrict__ pos, const float* __restrict__ scales,
                                                  const float* __restrict__ rot, int n, float sh0, float sh1,
                                                  float sh2, float* __restrict__ center, float* __restrict__ radius,
                                                  float* __restrict__ cov, float* __restrict__ half_shape) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g == 0) {
    half_shape[0] = sh0 * 0.5f;
    half_shape[1] = sh1 * 0.5f;
    half_shape[2] = sh2 * 0.5f;
  }
  if (g >= n) return;
  float c[3], M[9], rad;
  prologue_one(g, pos, scales, rot, sh0, sh1, sh2, center, radius, cov, c, M, rad);
}

// k_prologue and k_prep in one pass (the fused call): centre, radius and
// cov_inv written as k_prologue writes them, and the record (and pair count)
// built from the same fp32 values as k_prep builds it from memory.  With gacc,
// each Gaussian's moment row is zeroed too (the adjoint accumulates into it).
__global__ __launch_bounds__(256) void k_prologue_prep(const float* __restrict__ pos,
                                                       const float* __restrict__ scales,
                                                       const float* __restrict__ rot, int n, float sh0, float sh1,
                                                       float sh2, float* __restrict__ center,
                                                       float* __restrict__ radius, float* __restrict__ cov,
                                                       float* __restrict__ half_shape,
                                                       const float* __restrict__ rho_r,
                                                       const float* __restrict__ rho_i, int rho_stride,
                                                       GRec* __restrict__ rec, ZRec* __restrict__ zt, BinGeom geo,
                                                       int32_t* __restrict__ counts, float* __restrict__ gacc) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g == 0) {
    half_shape[0] = sh0 * 0.5f;
    half_shape[1] = sh1 * 0.5f;
    half_shape[2] = sh2 * 0.5f;
  }
  if (gacc != nullptr) {
    // The block's 256 moment rows are one contiguous range: zeroed with
    // coalesced 16-byte stores (consecutive lanes, consecutive addresses).
    constexpr int kVec = 256 * kAccStride / 4;
    float4* rows = reinterpret_cast<float4*>(gacc + (int64_t)blockIdx.x * 256 * kAccStride);
    const int valid = min(256, n - (int)blockIdx.x * 256) * (kAccStride / 4);
#pragma unroll
    for (int i = threadIdx.x; i < kVec; i += 256)
      if (i < valid) rows[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  if (counts != nullptr && g == n) counts[n] = 0;
  if (g >= n) return;
  float c[3], M[9], rad;
  prologue_one(g, pos, scales, rot, sh0, sh1, sh2, center, radius, cov, c, M, rad);
  prep_one(g, c, M, sh0 * 0.5f, sh1 * 0.5f, sh2 * 0.5f, rho_r, rho_i, rho_stride, rec, zt, rad, geo, counts);
}

// Chain rule of the prologue: from (dL/dc, dL/dM) to positions, linear scales
// and quaternions.  dL/dT = (G + G^T) T;  dL/dR_ik = dL/dT_ik * si_k;
// dL/dsi_k = sum_i dL/dT_ik R_ik;  dL/ds_k = -si_k^2 dL/dsi_k;  dL/dq from
// dL/dR through R(q^) and the normalisation q^ = q / (|q| + 1e-8).  The
// radius carries no gradient (the reference computes it under no_grad).
// One Gaussian's chain rule through the prologue, from its grad_center (gc)
// and grad_cov_inv (gm), in fp64 (N-sized work; the products of large
// 1/scale factors with gradient terms of both signs cancel for strongly
// anisotropic Gaussians).  The parameters are the fp32 values the forward
// read; the derivative is that of the prologue's mathematics at them.
__device__ __forceinline__ void epilogue_one(const float* __restrict__ scales, const float* __restrict__ rot, int g,
                                             const double gc[3], const double gm[9], float sh0, float sh1, float sh2,
                                             float* __restrict__ g_pos, float* __restrict__ g_scales,
                                             float* __restrict__ g_rot) {
  const int64_t g3 = 3 * (int64_t)g;
  g_pos[g3 + 0] = (float)(gc[0] * (0.5 * sh0));
  g_pos[g3 + 1] = (float)(gc[1] * (0.5 * sh1));
  g_pos[g3 + 2] = (float)(gc[2] * (0.5 * sh2));
  const double s[3] = {scales[g3 + 0], scales[g3 + 1], scales[g3 + 2]};
  const float4 q4 = *reinterpret_cast<const float4*>(rot + 4 * (int64_t)g);
  const double qw = q4.x, qx = q4.y, qy = q4.z, qz = q4.w;
  const double nr = sqrt(qw * qw + qx * qx + qy * qy + qz * qz);
  const double nrm = nr + 1e-8;
  const double inv = 1.0 / nrm;
  const double w = qw * inv, x = qx * inv, y = qy * inv, z = qz * inv;
  const double R[3][3] = {{1.0 - 2.0 * y * y - 2.0 * z * z, 2.0 * x * y - 2.0 * z * w, 2.0 * x * z + 2.0 * y * w},
                          {2.0 * x * y + 2.0 * z * w, 1.0 - 2.0 * x * x - 2.0 * z * z, 2.0 * y * z - 2.0 * x * w},
                          {2.0 * x * z - 2.0 * y * w, 2.0 * y * z + 2.0 * x * w, 1.0 - 2.0 * x * x - 2.0 * y * y}};
  const double si[3] = {1.0 / (s[0] + 1e-8), 1.0 / (s[1] + 1e-8), 1.0 / (s[2] + 1e-8)};
  double T[3][3], G[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      T[i][k] = R[i][k] * si[k];
      G[i][k] = gm[3 * i + k];
    }
  double GR[3][3], gsi[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      double gt = 0.0;
#pragma unroll
      for (int j = 0; j < 3; ++j) gt += (G[a][j] + G[j][a]) * T[j][b];
      GR[a][b] = gt * si[b];
      gsi[b] += gt * R[a][b];
    }
#pragma unroll
  for (int k = 0; k < 3; ++k) g_scales[g3 + k] = (float)(-si[k] * si[k] * gsi[k]);
  const double gw = 2.0 * (x * (GR[2][1] - GR[1][2]) + y * (GR[0][2] - GR[2][0]) + z * (GR[1][0] - GR[0][1]));
  const double gx = 2.0 * (y * (GR[0][1] + GR[1][0]) + z * (GR[0][2] + GR[2][0]) + w * (GR[2][1] - GR[1][2]) -
                           2.0 * x * (GR[1][1] + GR[2][2]));
  const double gy = 2.0 * (x * (GR[0][1] + GR[1][0]) + z * (GR[1][2] + GR[2][1]) + w * (GR[0][2] - GR[2][0]) -
                           2.0 * y * (GR[0][0] + GR[2][2]));
  const double gz = 2.0 * (x * (GR[0][2] + GR[2][0]) + y * (GR[1][2] + GR[2][1]) + w * (GR[1][0] - GR[0][1]) -
                           2.0 * z * (GR[0][0] + GR[1][1]));
  // Through q^ = q / (|q| + eps): dL/dq = g / (n+eps) - q (g . q) / (n (n+eps)^2); no second term at n = 0.
  const double dot = gw * qw + gx * qx + gy * qy + gz * qz;
  const double c2 = nr > 0.0 ? dot / (nr * nrm * nrm) : 0.0;
  float4 out;
  out.x = (float)(gw * inv - qw * c2);
  out.y = (float)(gx * inv - qx * c2);
  out.z = (float)(gy * inv - qy * c2);
  out.w = (float)(gz * inv - qz * c2);
  *reinterpret_cast<float4*>(g_rot + 4 * (int64_t)g) = out;
}


__global__ __launch_bounds__(256) void k_epilogue(const float* __restrict__ scales, const float* __restrict__ rot,
                                                  const float* __restrict__ g_center, const float* __restrict__ g_cov,
                                                  int n, float sh0, float sh1, float sh2, float* __restrict__ g_pos,
                                                  float* __restrict__ g_scales, float* __restrict__ g_rot) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= n) return;
  double gc[3], gm[9];
#pragma unroll
  for (int i = 0; i < 3; ++i) gc[i] = g_center[3 * (int64_t)g + i];
#pragma unroll
  for (int i = 0; i < 9; ++i) gm[i] = g_cov[9 * (int64_t)g + i];
  epilogue_one(scales, rot, g, gc, gm, sh0, sh1, sh2, g_pos, g_scales, g_rot);
}

// k_finalize and k_epilogue in one pass: the adjoint of the fused call from
// the moments straight to the parameters' gradients (grad_center and
// grad_cov_inv stay in registers).
__global__ __launch_bounds__(256) void k_finalize_epilogue(const float* __restrict__ gacc,
                                                           const float* __restrict__ cov,
                                                           const float* __restrict__ half_shape,
                                                           const float* __restrict__ scales,
                                                           const float* __restrict__ rot, int n, float sh0,
                                                           float sh1, float sh2, float* __restrict__ g_pos,
                                                           float* __restrict__ g_scales, float* __restrict__ g_rot,
                                                           float* __restrict__ g_rr, float* __restrict__ g_ri,
                                                           int rho_stride) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= n) return;
  const float hs[3] = {half_shape[0], half_shape[1], half_shape[2]};
  double gc[3], gm[9];
  float grr, gri;
  finalize_one(gacc, cov, hs, g, gc, gm, grr, gri);
  g_rr[(int64_t)g * rho_stride] = grr;
  g_ri[(int64_t)g * rho_stride] = gri;
  epilogue_one(scales, rot, g, gc, gm, sh0, sh1, sh2, g_pos, g_scales, g_rot);
}

// Evaluations the backward makes for one pair (record r, tile of b): over
// the live columns of its box -- the 4x4 (H, W) column blocks with the D rows
// their column bound leaves, as k_backward stages them -- rows x the column's
// valid H x W voxels.  nb: column blocks per tile axis.
__device__ __forceinline__ unsigned int bwd_pair_evals(const GRec& r, const ZCoefs& zc, const BlockCtx& b, int nb) {
  const float lo0 = fmaxf(ceilf(r.c0 - r.e0), (float)b.o0);
  const float lo1 = fmaxf(ceilf(r.c1 - r.e1), (float)b.o1);
  const float lo2 = fmaxf(ceilf(r.c2 - r.e2), (float)b.o2);
  const float hi0 = fminf(floorf(r.c0 + r.e0), (float)(b.o0 + b.e0 - 1));
  const float hi1 = fminf(floorf(r.c1 + r.e1), (float)(b.o1 + b.e1 - 1));
  const float hi2 = fminf(floorf(r.c2 + r.e2), (float)(b.o2 + b.e2 - 1));
  if (!(lo0 <= hi0 && lo1 <= hi1 && lo2 <= hi2)) return 0u;
  const int y_lo = (int)lo1 - b.o1, y_hi = (int)hi1 - b.o1;
  const int x_lo = (int)lo2 - b.o2, x_hi = (int)hi2 - b.o2;
  const int l1 = y_lo >> 2, h1 = y_hi >> 2, l2 = x_lo >> 2, h2 = x_hi >> 2;
  float hmn[kMaxBlk], hmx[kMaxBlk], wmn[kMaxBlk], wmx[kMaxBlk];
#pragma unroll
  for (int j = 0; j < kMaxBlk; ++j) {
    hmn[j] = wmn[j] = -INFINITY;
    hmx[j] = wmx[j] = INFINITY;
    if (zc.ok) {
      const int y0 = max(4 * j, y_lo), y1 = min(4 * j + 3, y_hi);
      const int x0 = max(4 * j, x_lo), x1 = min(4 * j + 3, x_hi);
      z_span(zc.h, (float)(b.o1 + y0) - r.c1, (float)(b.o1 + y1) - r.c1, hmn[j], hmx[j]);
      z_span(zc.w, (float)(b.o2 + x0) - r.c2, (float)(b.o2 + x1) - r.c2, wmn[j], wmx[j]);
    }
  }
  unsigned int total = 0;
#pragma unroll
  for (int j1 = 0; j1 < kMaxBlk; ++j1) {
#pragma unroll
    for (int j2 = 0; j2 < kMaxBlk; ++j2) {
      if (j1 < nb && j2 < nb && l1 <= j1 && j1 <= h1 && l2 <= j2 && j2 <= h2) {
        float zl, zh;
        column_rows(r.c0, hmn[j1], hmx[j1], wmn[j2], wmx[j2], lo0, hi0, zl, zh);
        if (zl <= zh) {
          const int rows = (int)zh - (int)zl + 1;
          const int nh = min(kSub, b.e1 - kSub * j1), nw = min(kSub, b.e2 - kSub * j2);
          total += (unsigned int)(rows * nh * nw);
        }
      }
    }
  }
  return total;
}

// Exact evaluation counts for the roofline accounting: the forward kernel's
// loop structure with the accumulation replaced by counting.
__global__ __launch_bounds__(kTileThreads) void k_count(const GRec* __restrict__ rec,
                                                        const ZRec* __restrict__ zt,
                                                        const int32_t* __restrict__ ranges,
                                                        const int32_t* __restrict__ pair_ids, VoxGeom geo,
                                                        unsigned long long* __restrict__ counts) {
  __shared__ Stage sh;
  const int logical = xcd_remap(blockIdx.x, gridDim.x);
  if (logical >= geo.blocks) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int gl = lane & 15, xi = lane >> 4;
  const BlockCtx b = make_ctx(logical, geo, ranges, wave);
  if (!b.valid) return;
  const bool x_ok = xi < b.n2;
  if (threadIdx.x == 0 && logical % geo.chunks == 0) {
    const unsigned long long pairs = (unsigned long long)(b.p_end - b.p_begin);
    atomicAdd(&counts[0], pairs * (unsigned long long)(b.e0 * b.e1 * b.e2));
  }
  unsigned int active = 0, culled = 0;
  unsigned long long bwd = 0;
  const bool use_mask = geo.nsb <= 2;
  const bool first_chunk = logical % geo.chunks == 0;  // the backward's evaluations are counted once per tile
  if (threadIdx.x == 0) sh.n_touch = 0;
  for (int bs = b.p_begin; bs < b.p_end; bs += kBatch) {
    const int cnt = min(kBatch, b.p_end - bs);
    __syncthreads();
    bool touch = false;
    if ((int)threadIdx.x < cnt) {
      const int gid = pair_ids[bs + threadIdx.x];
      const GRec r = rec[gid];
      touch = stage_pair(sh, threadIdx.x, r, gid, b);
      if (first_chunk && touch) bwd += bwd_pair_evals(r, z_load(zt, gid), b, geo.nsb);
    }
    if (use_mask) {
      append_touch(sh, touch, threadIdx.x, lane);
      __syncthreads();
      batch_masks(sh, b, zt);
    }
    __syncthreads();
    if (threadIdx.x == 0) sh.n_touch = 0;
    if (!b.active) continue;
    const int nhit = build_list(sh, cnt, b, wave, lane, use_mask);
    for (int gb = 0; gb < nhit; gb += 16) {
      const int idx = gb + gl;
      const bool g_ok = (idx < nhit) && x_ok;
      const int k = sh.list[wave][idx < nhit ? idx : gb];
      const GaussLane g = lane_terms(load_raw(sh, k), b, xi);
#pragma unroll
      for (int z = 0; z < kSub; ++z) {
        if (z < b.n0) {
          const float u0 = (float)(b.o0 + b.s0 + z) - g.c0;
          float t1, r;
          row_terms(g, u0, t1, r);
#pragma unroll
          for (int y = 0; y < kSub; ++y) {
            if (y < b.n1) {
              const float y1 = fmaf(g.y1h, g.u1[y], t1);
              const float q = fmaf(-y1, y1, r);
              active += (g_ok && (q >= kThr)) ? 1u : 0u;
              culled += g_ok ? 1u : 0u;
            }
          }
        }
      }
    }
  }
  unsigned long long a = active, c = culled;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off, 64);
    c += __shfl_xor(c, off, 64);
    bwd += __shfl_xor(bwd, off, 64);
  }
  if (lane == 0) {
    atomicAdd(&counts[1], a);
    atomicAdd(&counts[2], c);
    atomicAdd(&counts[3], bwd);
  }
}

// ------------------------------------------------------------- host helpers --

// Any positive tile edge (the reference takes any, gaussian/tile_voxelizer.py:248-251);
// the tile kernels split tiles above 16 into sub-tiles (VoxGeom).
bool tile_geometry(int32_t d, int32_t h, int32_t w, int32_t tile, int nt[3], int64_t* num_tiles) {
  if (d <= 0 || h <= 0 || w <= 0 || tile <= 0) return false;
  nt[0] = ceil_div_i(d, tile);
  nt[1] = ceil_div_i(h, tile);
  nt[2] = ceil_div_i(w, tile);
  *num_tiles = (int64_t)nt[0] * nt[1] * nt[2];
  return *num_tiles < (int64_t)INT_MAX;
}

int radix_passes(int num_tiles) {
  int bits = 0;
  while ((1ll << bits) < num_tiles) ++bits;
  return (bits + kRadixBits - 1) / kRadixBits;
}

int key_bits(int num_tiles) {
  int bits = 0;
  while ((1ll << bits) < num_tiles) ++bits;
  return bits;
}

VoxGeom make_vox_geom(int d, int h, int w, int tile, const int nt[3], int num_tiles) {
  VoxGeom g;
  g.D = d;
  g.H = h;
  g.W = w;
  g.tile = tile;
  g.nTD = nt[0];
  g.nTH = nt[1];
  g.nTW = nt[2];
  g.num_tiles = num_tiles;
  g.parts = ceil_div_i(tile, kMaxSubTile);
  g.sub = ceil_div_i(tile, g.parts);
  g.nSD = nt[0] * g.parts;
  g.nSH = nt[1] * g.parts;
  g.nSW = nt[2] * g.parts;
  g.num_sub = g.nSD * g.nSH * g.nSW;
  g.nsb = ceil_div_i(g.sub, kSub);
  g.chunks = ceil_div_i(g.nsb * g.nsb * g.nsb, kWaves);
  g.blocks = g.num_sub * g.chunks;
  return g;
}

// Whether the tile kernels' grid of sub-tiles fits int indexing.
bool vox_geom_fits(const int nt[3], int tile) {
  const int64_t parts = ceil_div_i(tile, kMaxSubTile);
  const int64_t sub = ceil_div_i(tile, (int)parts);
  const int64_t nsb = (sub + kSub - 1) / kSub;
  const int64_t chunks = (nsb * nsb * nsb + kWaves - 1) / kWaves;
  const int64_t blocks = nt[0] * parts * nt[1] * parts * nt[2] * parts * chunks;
  return nt[0] * parts < INT_MAX && nt[1] * parts < INT_MAX && nt[2] * parts < INT_MAX && blocks < INT_MAX - 8;
}

// Binning workspace: one key/value buffer pair (the first radix pass output,
// or the expansion when there is a single tile), a second pair only when three
// or more radix passes ping-pong, the digit histograms and scan sums, and the
// per-block first Gaussians.
size_t fill_ws_layout(int64_t total, int num_tiles, Carve& cv, int32_t** ka, int32_t** va, int32_t** kb,
                      int32_t** vb, int32_t** hist, int64_t** sums, int32_t** first) {
  const int64_t nblk = ceil_div_l(total, kRadixTile);
  const int64_t hist_n = (int64_t)(1 << kRadixBits) * nblk;
  const size_t second = radix_passes(num_tiles) >= 3 ? (size_t)total : 0;
  *ka = cv.take<int32_t>((size_t)total);
  *va = cv.take<int32_t>((size_t)total);
  *kb = cv.take<int32_t>(second);
  *vb = cv.take<int32_t>(second);
  *hist = cv.take<int32_t>((size_t)hist_n);
  *sums = cv.take<int64_t>((size_t)ceil_div_l(hist_n, kScanTile) + 1);
  *first = cv.take<int32_t>((size_t)nblk + 1);
  return align_up(cv.off, 256);
}

// Forward workspace: the records and the column-bound records.
size_t fwd_layout(void* base, int32_t n, GRec** rec, ZRec** zt) {
  Carve cv(base);
  GRec* r = cv.take<GRec>((size_t)n);
  ZRec* z = cv.take<ZRec>((size_t)n);
  if (rec != nullptr) *rec = r;
  if (zt != nullptr) *zt = z;
  return align_up(cv.off, 256);
}

// Opt-in kernel timing (gstv_profile_begin / gstv_profile_end).
constexpr int kProfMax = 1024;
struct Profile {
  bool on = false;
  hipEvent_t start[2][kProfMax] = {};
  hipEvent_t stop[2][kProfMax] = {};
  int used[2] = {0, 0};
};
Profile g_prof;

// Records the start event of a kernel launch of `kind` (0 forward, 1 backward);
// returns the slot to close with prof_stop, or -1 when not recording.
int prof_start(int kind, hipStream_t s) {
  if (!g_prof.on || g_prof.used[kind] >= kProfMax) return -1;
  const int i = g_prof.used[kind];
  if (g_prof.start[kind][i] == nullptr) {
    if (hipEventCreate(&g_prof.start[kind][i]) != hipSuccess) return -1;
    if (hipEventCreate(&g_prof.stop[kind][i]) != hipSuccess) return -1;
  }
  if (hipEventRecord(g_prof.start[kind][i], s) != hipSuccess) return -1;
  return i;
}

void prof_stop(int kind, int slot, hipStream_t s) {
  if (slot < 0) return;
  if (hipEventRecord(g_prof.stop[kind][slot], s) == hipSuccess) g_prof.used[kind] = slot + 1;
}


// ---------------------------------------------- no-allocation workspace --
//
// gstv_backward_noalloc keeps its workspace (the keep layout: records,
// column-bound records, moment accumulator, scan sums) in this file-scope
// device array instead of a caller's buffer.  Every GPU the library runs on
// holds its own copy, at its own address, until the process ends; in host
// code `scratch` names a host-side stand-in, so kernels get the device address
// from hipGetSymbolAddress, resolved once per device with that device current.
constexpr size_t kNoAllocScratchBytes =
    (size_t)GSTV_NOALLOC_MAX_GAUSSIANS * (sizeof(GRec) + sizeof(ZRec) + kAccStride * sizeof(float) + sizeof(uint32_t)) +
    (1u << 16);
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

// Device address of the current device's scratch array.
int scratch_of_current_device(void** out) {
  int dev = 0;
  int st = hip_status(hipGetDevice(&dev), "noalloc_get_device");
  if (st != GSTV_OK) return st;
  if (dev < 0 || dev >= kMaxDevices) return GSTV_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    st = hip_status(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "noalloc_symbol_address");
    if (st != GSTV_OK) return st;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return GSTV_OK;
}

}  // namespace

// ====================================================================== C ABI

extern "C" {

const char* gstv_status_string(int status) {
  switch (status) {
    case GSTV_OK: return "success";
    case GSTV_ERR_INVALID_ARGUMENT: return "invalid argument";
    case GSTV_ERR_WORKSPACE_TOO_SMALL: return "workspace too small";
    case GSTV_ERR_TOO_MANY_PAIRS: return "too many (tile, Gaussian) pairs: total must be < 2^31";
    case GSTV_ERR_HIP: return g_last_detail[0] ? g_last_detail : "HIP runtime error";
    case GSTV_ERR_TOO_MANY_GAUSSIANS:
      return "too many Gaussians for gstv_backward_noalloc (at most GSTV_NOALLOC_MAX_GAUSSIANS = 262144)";
    default: return "unknown status";
  }
}

int gstv_version(void) { return 101; }

size_t gstv_bin_count_workspace_bytes(int32_t n) {
  if (n < 0) return 0;
  Carve cv(nullptr);
  cv.take<int64_t>((size_t)ceil_div_l((int64_t)n + 1, kScanTile) + 1);
  return align_up(cv.off, 256);
}

namespace {
int bin_count_impl(const GRec* rec, const float* center_vox, const float* radius_vox, int32_t n, int32_t d,
                   int32_t h, int32_t w, int32_t tile, int32_t* gauss_offsets, int64_t* total_pairs,
                   void* workspace, size_t workspace_bytes, void* stream) {
  BinGeom geo;
  geo.rec = rec;
  geo.live = nullptr;
  int64_t num_tiles;
  if (n < 0 || gauss_offsets == nullptr || total_pairs == nullptr ||
      !tile_geometry(d, h, w, tile, geo.nt, &num_tiles))
    return GSTV_ERR_INVALID_ARGUMENT;
  if (n > 0 && (center_vox == nullptr || radius_vox == nullptr)) return GSTV_ERR_INVALID_ARGUMENT;
  if (workspace_bytes < gstv_bin_count_workspace_bytes(n) || workspace == nullptr)
    return GSTV_ERR_WORKSPACE_TOO_SMALL;
  geo.size[0] = d;
  geo.size[1] = h;
  geo.size[2] = w;
  geo.tile = tile;
  geo.tile_shift = tile_shift_of(tile);
  hipStream_t s = static_cast<hipStream_t>(stream);
  Carve cv(workspace);
  int64_t* sums = cv.take<int64_t>((size_t)ceil_div_l((int64_t)n + 1, kScanTile) + 1);
  hipLaunchKernelGGL(k_bin_count, dim3(ceil_div_i(n + 1, 256)), dim3(256), 0, s, center_vox, radius_vox, n, geo,
                     gauss_offsets);
  GSTV_LAUNCH_CHECK("bin_count");
  return launch_exclusive_scan(gauss_offsets, (int64_t)n + 1, gauss_offsets, sums, total_pairs, s);
}
}  // namespace

int gstv_bin_count(const float* center_vox, const float* radius_vox, int32_t n, int32_t d, int32_t h,
                   int32_t w, int32_t tile, int32_t* gauss_offsets, int64_t* total_pairs, void* workspace,
                   size_t workspace_bytes, void* stream) {
  return bin_count_impl(nullptr, center_vox, radius_vox, n, d, h, w, tile, gauss_offsets, total_pairs, workspace,
                        workspace_bytes, stream);
}

size_t gstv_bin_fill_workspace_bytes(int32_t n, int64_t total_pairs, int32_t num_tiles) {
  (void)n;
  if (total_pairs < 0) return 0;
  Carve cv(nullptr);
  int32_t *ka, *va, *kb, *vb, *hist, *first;
  int64_t* sums;
  return fill_ws_layout(total_pairs, num_tiles, cv, &ka, &va, &kb, &vb, &hist, &sums, &first);
}

namespace {
// total_dev null: total_pairs is the pair total.  Otherwise total_pairs is a
// capacity and the kernels read the total from *total_dev (see PairTotal).
int bin_fill_impl(const GRec* rec, uint32_t* live, const float* center_vox, const float* radius_vox, int32_t n,
                  int32_t d, int32_t h, int32_t w, int32_t tile, const int32_t* gauss_offsets, int64_t total_pairs,
                  const int64_t* total_dev, int32_t* tile_ranges, int32_t* pair_ids, void* workspace,
                  size_t workspace_bytes, void* stream) {
  BinGeom geo;
  geo.rec = rec;
  geo.live = live;
  int64_t num_tiles64;
  if (n < 0 || total_pairs < 0 || tile_ranges == nullptr || !tile_geometry(d, h, w, tile, geo.nt, &num_tiles64))
    return GSTV_ERR_INVALID_ARGUMENT;
  if (total_pairs >= (int64_t)INT_MAX) return GSTV_ERR_TOO_MANY_PAIRS;
  const int num_tiles = (int)num_tiles64;
  geo.size[0] = d;
  geo.size[1] = h;
  geo.size[2] = w;
  geo.tile = tile;
  geo.tile_shift = tile_shift_of(tile);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int total = (int)total_pairs;
  const PairTotal pt{total_dev, total};
  const int passes = radix_passes(num_tiles);
  if (total == 0 || passes == 0) {
    hipLaunchKernelGGL(k_fill_i32, dim3(ceil_div_i(2 * num_tiles, 256)), dim3(256), 0, s, tile_ranges,
                       (int64_t)2 * num_tiles, 0);
    GSTV_LAUNCH_CHECK("ranges_clear");
    if (total == 0) return GSTV_OK;
  }
  if (n <= 0 || center_vox == nullptr || radius_vox == nullptr || gauss_offsets == nullptr ||
      pair_ids == nullptr)
    return GSTV_ERR_INVALID_ARGUMENT;
  if (workspace == nullptr || workspace_bytes < gstv_bin_fill_workspace_bytes(n, total_pairs, num_tiles))
    return GSTV_ERR_WORKSPACE_TOO_SMALL;
  Carve cv(workspace);
  int32_t *ka, *va, *kb, *vb, *hist, *first;
  int64_t* sums;
  fill_ws_layout(total_pairs, num_tiles, cv, &ka, &va, &kb, &vb, &hist, &sums, &first);

  if (passes == 0) {
    // A single tile: the expansion order is final.
    hipLaunchKernelGGL(k_bin_expand, dim3(ceil_div_i(total, kExpandTile)), dim3(kExpandThreads), 0, s, center_vox,
                       radius_vox, n, geo, gauss_offsets, pt, ka, pair_ids);
    GSTV_LAUNCH_CHECK("bin_expand");
    hipLaunchKernelGGL(k_ranges, dim3(ceil_div_i(total, 256)), dim3(256), 0, s, ka, pt, tile_ranges);
    GSTV_LAUNCH_CHECK("ranges");
    return GSTV_OK;
  }

  hipLaunchKernelGGL(k_block_first, dim3(ceil_div_i(std::max(n, num_tiles), 256)), dim3(256), 0, s, gauss_offsets, n,
                     pt, first, tile_ranges, num_tiles);
  GSTV_LAUNCH_CHECK("block_first");
  const int bits_total = key_bits(num_tiles);
  const int width = (bits_total + passes - 1) / passes;  // balanced digit widths
  const int nblk = ceil_div_i(total, kRadixTile);
  int32_t* buf_k[2] = {ka, kb};
  int32_t* buf_v[2] = {va, vb};
  int cur = -1;  // -1: the first pass generates its pairs
  for (int p = 0; p < passes; ++p) {
    const int shift = p * width;
    const int bits = std::min(width, bits_total - shift);
    const bool last = p == passes - 1;
    PairSource src;
    src.keys = cur >= 0 ? buf_k[cur] : nullptr;
    src.vals = cur >= 0 ? buf_v[cur] : nullptr;
    src.center = center_vox;
    src.radius = radius_vox;
    src.offsets = gauss_offsets;
    src.first = first;
    src.n = n;
    src.geo = geo;
    const int dst = cur >= 0 ? 1 - cur : 0;
    int32_t* kout = buf_k[dst];
    int32_t* vout = last ? pair_ids : buf_v[dst];
    if (cur < 0)
      hipLaunchKernelGGL(k_radix_hist<true>, dim3(nblk), dim3(kRadixThreads), 0, s, src, pt, shift, bits, nblk,
                         hist);
    else
      hipLaunchKernelGGL(k_radix_hist<false>, dim3(nblk), dim3(kRadixThreads), 0, s, src, pt, shift, bits, nblk,
                         hist);
    GSTV_LAUNCH_CHECK("radix_hist");
    const int64_t hist_n = (int64_t)(1 << bits) * nblk;
    int st = launch_exclusive_scan(hist, hist_n, hist, sums, nullptr, s);
    if (st != GSTV_OK) return st;
    if (cur < 0 && last)
      hipLaunchKernelGGL((k_radix_scatter_gen<true>), dim3(nblk), dim3(kRadixThreads), 0, s, src, pt, shift,
                         bits, nblk, hist, kout, vout, tile_ranges);
    else if (cur < 0)
      hipLaunchKernelGGL((k_radix_scatter_gen<false>), dim3(nblk), dim3(kRadixThreads), 0, s, src, pt, shift,
                         bits, nblk, hist, kout, vout, tile_ranges);
    else if (last)
      hipLaunchKernelGGL((k_radix_scatter<true>), dim3(nblk), dim3(kRadixThreads), 0, s, src, pt, shift,
                         bits, nblk, hist, kout, vout, tile_ranges);
    else
      hipLaunchKernelGGL((k_radix_scatter<false>), dim3(nblk), dim3(kRadixThreads), 0, s, src, pt, shift,
                         bits, nblk, hist, kout, vout, tile_ranges);
    GSTV_LAUNCH_CHECK("radix_scatter");
    cur = dst;
  }
  hipLaunchKernelGGL(k_ranges_fix, dim3(ceil_div_i(num_tiles, 256)), dim3(256), 0, s, tile_ranges, num_tiles);
  GSTV_LAUNCH_CHECK("ranges_fix");
  return GSTV_OK;
}
}  // namespace

int gstv_bin_fill(const float* center_vox, const float* radius_vox, int32_t n, int32_t d, int32_t h,
                  int32_t w, int32_t tile, const int32_t* gauss_offsets, int64_t total_pairs,
                  int32_t* tile_ranges, int32_t* pair_ids, void* workspace, size_t workspace_bytes,
                  void* stream) {
  return bin_fill_impl(nullptr, nullptr, center_vox, radius_vox, n, d, h, w, tile, gauss_offsets, total_pairs, nullptr,
                       tile_ranges, pair_ids, workspace, workspace_bytes, stream);
}

size_t gstv_forward_workspace_bytes(int32_t n) {
  if (n < 0) return 0;
  return fwd_layout(nullptr, n, nullptr, nullptr);
}

namespace {
int check_tile_args(const int32_t* tile_ranges, int32_t num_tiles, int32_t n, int32_t tile, int32_t d, int32_t h,
                    int32_t w, int nt[3]) {
  int64_t expect;
  if (n < 0 || !tile_geometry(d, h, w, tile, nt, &expect) || !vox_geom_fits(nt, tile))
    return GSTV_ERR_INVALID_ARGUMENT;
  if ((int64_t)num_tiles != expect || tile_ranges == nullptr) return GSTV_ERR_INVALID_ARGUMENT;
  return GSTV_OK;
}

int launch_prep(const float* center_vox, const float* cov_inv, const float* rho_r, const float* rho_i,
                int rho_stride, const float* half_shape, int32_t n, GRec* rec, ZRec* zt, hipStream_t s,
                const float* radius = nullptr, const BinGeom* geo = nullptr, int32_t* counts = nullptr) {
  if (n == 0 && counts == nullptr) return GSTV_OK;
  BinGeom g0{};
  hipLaunchKernelGGL(k_prep, dim3(ceil_div_i(counts != nullptr ? n + 1 : n, 256)), dim3(256), 0, s, center_vox,
                     cov_inv, rho_r, rho_i, rho_stride, half_shape, n, rec, zt, radius, geo != nullptr ? *geo : g0,
                     counts);
  GSTV_LAUNCH_CHECK("prep");
  return GSTV_OK;
}

int tile_grid(const VoxGeom& g) { return ((g.blocks + 7) / 8) * 8; }
}  // namespace

namespace {
// Forward tile kernel on built records.  pair_columns: written when non-null
// and the tile is at most 8 voxels per axis.
void launch_forward(const GRec* rec, const ZRec* zt, const int32_t* tile_ranges, int32_t num_tiles,
                    const int32_t* pair_ids, int32_t tile, int32_t d, int32_t h, int32_t w, const int nt[3],
                    float* vol_r, float* vol_i, uint32_t* pair_columns, hipStream_t s) {
  const VoxGeom geo = make_vox_geom(d, h, w, tile, nt, num_tiles);
  const int slot = prof_start(0, s);
  hipLaunchKernelGGL(k_forward, dim3(tile_grid(geo)), dim3(kTileThreads), 0, s, rec, zt, tile_ranges, pair_ids, geo,
                     vol_r, vol_i, geo.nsb <= 2 ? pair_columns : nullptr);
  prof_stop(0, slot, s);
}

void launch_clear_acc(float* gacc, int32_t n, hipStream_t s) {
  const int64_t nacc = (int64_t)n * kAccStride;
  hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)ceil_div_l(nacc, 256)), dim3(256), 0, s, gacc, nacc, 0.0f);
}

// Backward tile kernel into a zeroed moment accumulator, then (when
// grad_center is given) the moments -> gradients pass.  pair_columns: the
// forward's column words (tiles of at most 8 voxels per axis), or null to
// compute the column bounds here.
void launch_backward(const GRec* rec, const ZRec* zt, float* gacc, const float* cov_inv, const float* half_shape,
                     const int32_t* tile_ranges, int32_t num_tiles, const int32_t* pair_ids, int32_t n,
                     const float* grad_r, const float* grad_i, int32_t tile, int32_t d, int32_t h, int32_t w,
                     const int nt[3], float* grad_center, float* grad_cov_inv, float* grad_rho_r, float* grad_rho_i,
                     int rho_stride, const uint32_t* pair_columns, hipStream_t s) {
  const VoxGeom geo = make_vox_geom(d, h, w, tile, nt, num_tiles);
  const int slot = prof_start(1, s);
  VoxGeom geo_b = geo;  // the backward runs one block per tile
  geo_b.chunks = 1;
  geo_b.blocks = geo.num_sub;
  const int tp = geo.nsb * kSub;
  // 8-voxel tiles: 8 zero slabs past the tile (k_backward<8>).
  const size_t lds_grad = (size_t)(geo.sub == 8 ? 2 * tp : tp) * grad_zs(tp) * sizeof(float2);
  const uint32_t* cols = geo.nsb <= 2 ? pair_columns : nullptr;
  if (geo.sub == 8)
    hipLaunchKernelGGL(k_backward<8>, dim3(tile_grid(geo_b)), dim3(kTileThreads), lds_grad, s, rec, zt,
                       tile_ranges, pair_ids, geo_b, grad_r, grad_i, gacc, cols);
  else
    hipLaunchKernelGGL(k_backward<0>, dim3(tile_grid(geo_b)), dim3(kTileThreads), lds_grad, s, rec, zt,
                       tile_ranges, pair_ids, geo_b, grad_r, grad_i, gacc, cols);
  prof_stop(1, slot, s);
  if (grad_center != nullptr)
    hipLaunchKernelGGL(k_finalize, dim3(ceil_div_i(n, 256)), dim3(256), 0, s, (const float*)gacc, cov_inv,
                       half_shape, n, grad_center, grad_cov_inv, grad_rho_r, grad_rho_i, rho_stride);
}

// Keep workspace: the records, the column-bound records, then the 64-byte
// moment rows.
struct KeepLayout {
  GRec* rec;
  ZRec* zt;
  float* gacc;
  int64_t* sums;   // block sums of the pair-count scan
  uint32_t* live;  // live-tile masks of the binning (BinGeom::live)
};
size_t keep_layout(void* base, int32_t n, KeepLayout* out) {
  Carve cv(base);
  GRec* rec = cv.take<GRec>((size_t)n);
  ZRec* zt = cv.take<ZRec>((size_t)n);
  float* gacc = cv.take<float>((size_t)n * kAccStride);
  int64_t* sums = cv.take<int64_t>((size_t)ceil_div_l((int64_t)n + 1, kScanTile) + 1);
  uint32_t* live = cv.take<uint32_t>((size_t)n);
  if (out != nullptr) *out = KeepLayout{rec, zt, gacc, sums, live};
  return align_up(cv.off, 256);
}

}  // namespace

int gstv_forward(const float* center_vox, const float* cov_inv, const float* rho_r, const float* rho_i,
                 const int32_t* tile_ranges, int32_t num_tiles, const int32_t* pair_ids, const float* half_shape,
                 int32_t n, int32_t tile, int32_t d, int32_t h, int32_t w, float* vol_r, float* vol_i,
                 void* workspace, size_t workspace_bytes, void* stream) {
  int nt[3];
  int st = check_tile_args(tile_ranges, num_tiles, n, tile, d, h, w, nt);
  if (st != GSTV_OK) return st;
  if (vol_r == nullptr || vol_i == nullptr || half_shape == nullptr) return GSTV_ERR_INVALID_ARGUMENT;
  if (n > 0 && (center_vox == nullptr || cov_inv == nullptr || rho_r == nullptr || rho_i == nullptr))
    return GSTV_ERR_INVALID_ARGUMENT;
  if (n > 0 && (workspace == nullptr || workspace_bytes < gstv_forward_workspace_bytes(n)))
    return GSTV_ERR_WORKSPACE_TOO_SMALL;
  hipStream_t s = static_cast<hipStream_t>(stream);
  GRec* rec;
  ZRec* zt;
  fwd_layout(workspace, n, &rec, &zt);
  st = launch_prep(center_vox, cov_inv, rho_r, rho_i, 1, half_shape, n, rec, zt, s);
  if (st != GSTV_OK) return st;
  launch_forward(rec, zt, tile_ranges, num_tiles, pair_ids, tile, d, h, w, nt, vol_r, vol_i, nullptr, s);
  GSTV_LAUNCH_CHECK("forward");
  return GSTV_OK;
}

size_t gstv_backward_workspace_bytes(int32_t n) {
  if (n < 0) return 0;
  return keep_layout(nullptr, n, nullptr);
}

int gstv_backward(const float* center_vox, const float* cov_inv, const float* rho_r, const float* rho_i,
                  const int32_t* tile_ranges, int32_t num_tiles, const int32_t* pair_ids, const float* half_shape,
                  int32_t n, const float* grad_r, const float* grad_i, int32_t tile, int32_t d, int32_t h,
                  int32_t w, float* grad_center, float* grad_cov_inv, float* grad_rho_r, float* grad_rho_i,
                  void* workspace, size_t workspace_bytes, void* stream) {
  int nt[3];
  int st = check_tile_args(tile_ranges, num_tiles, n, tile, d, h, w, nt);
  if (st != GSTV_OK) return st;
  if (n == 0) return GSTV_OK;
  if (center_vox == nullptr || cov_inv == nullptr || rho_r == nullptr || rho_i == nullptr ||
      half_shape == nullptr || grad_r == nullptr || grad_i == nullptr || grad_center == nullptr ||
      grad_cov_inv == nullptr || grad_rho_r == nullptr || grad_rho_i == nullptr)
    return GSTV_ERR_INVALID_ARGUMENT;
  if (workspace == nullptr || workspace_bytes < gstv_backward_workspace_bytes(n))
    return GSTV_ERR_WORKSPACE_TOO_SMALL;
  hipStream_t s = static_cast<hipStream_t>(stream);
  KeepLayout ws;
  keep_layout(workspace, n, &ws);
  st = launch_prep(center_vox, cov_inv, rho_r, rho_i, 1, half_shape, n, ws.rec, ws.zt, s);
  if (st != GSTV_OK) return st;
  launch_clear_acc(ws.gacc, n, s);
  launch_backward(ws.rec, ws.zt, ws.gacc, cov_inv, half_shape, tile_ranges, num_tiles, pair_ids, n, grad_r, grad_i, tile, d,
                  h, w, nt, grad_center, grad_cov_inv, grad_rho_r, grad_rho_i, 1, nullptr, s);
  GSTV_LAUNCH_CHECK("backward");
  return GSTV_OK;
}

int gstv_backward_noalloc(const float* center_vox, const float* cov_inv, const float* rho_r, const float* rho_i,
                          const int32_t* tile_ranges, int32_t num_tiles, const int32_t* pair_ids,
                          const float* half_shape, int32_t n, const float* grad_r, const float* grad_i, int32_t tile,
                          int32_t d, int32_t h, int32_t w, float* grad_center, float* grad_cov_inv,
                          float* grad_rho_r, float* grad_rho_i, void* stream) {
  int nt[3];
  int st = check_tile_args(tile_ranges, num_tiles, n, tile, d, h, w, nt);
  if (st != GSTV_OK) return st;
  if (n > GSTV_NOALLOC_MAX_GAUSSIANS) return GSTV_ERR_TOO_MANY_GAUSSIANS;
  if (n == 0) return GSTV_OK;
  if (center_vox == nullptr || cov_inv == nullptr || rho_r == nullptr || rho_i == nullptr ||
      half_shape == nullptr || grad_r == nullptr || grad_i == nullptr || grad_center == nullptr ||
      grad_cov_inv == nullptr || grad_rho_r == nullptr || grad_rho_i == nullptr)
    return GSTV_ERR_INVALID_ARGUMENT;
  if (keep_layout(nullptr, n, nullptr) > kNoAllocScratchBytes) return GSTV_ERR_TOO_MANY_GAUSSIANS;
  void* ws_ptr = nullptr;
  st = scratch_of_current_device(&ws_ptr);
  if (st != GSTV_OK) return st;
  hipStream_t s = static_cast<hipStream_t>(stream);
  KeepLayout ws;
  keep_layout(ws_ptr, n, &ws);
  // Every scratch region is rewritten before it is read: records by the
  // prep kernel, the accumulator by a fill kernel (no memset node, so the
  // calls can be captured into a graph).
  st = launch_prep(center_vox, cov_inv, rho_r, rho_i, 1, half_shape, n, ws.rec, ws.zt, s);
  if (st != GSTV_OK) return st;
  launch_clear_acc(ws.gacc, n, s);
  launch_backward(ws.rec, ws.zt, ws.gacc, cov_inv, half_shape, tile_ranges, num_tiles, pair_ids, n, grad_r, grad_i,
                  tile, d, h, w, nt, grad_center, grad_cov_inv, grad_rho_r, grad_rho_i, 1, nullptr, s);
  GSTV_LAUNCH_CHECK("backward_noalloc");
  return GSTV_OK;
}

size_t gstv_keep_workspace_bytes(int32_t n) {
  if (n < 0) return 0;
  return keep_layout(nullptr, n, nullptr);
}

int gstv_keep_prepare(const float* center_vox, const float* radius_vox, const float* cov_inv, const float* rho_r,
                      const float* rho_i, int32_t rho_stride, const float* half_shape, int32_t n, int32_t d,
                      int32_t h, int32_t w, int32_t tile, int32_t* gauss_offsets, int64_t* total_pairs,
                      void* keep_workspace, size_t keep_workspace_bytes, void* stream) {
  if (n < 0 || rho_stride < 1 || ((gauss_offsets == nullptr) != (total_pairs == nullptr)))
    return GSTV_ERR_INVALID_ARGUMENT;
  const bool count = gauss_offsets != nullptr;
  BinGeom geo{};
  if (count) {
    int64_t num_tiles;
    if (!tile_geometry(d, h, w, tile, geo.nt, &num_tiles)) return GSTV_ERR_INVALID_ARGUMENT;
    geo.size[0] = d;
    geo.size[1] = h;
    geo.size[2] = w;
    geo.tile = tile;
    geo.tile_shift = tile_shift_of(tile);
    if (n > 0 && radius_vox == nullptr) return GSTV_ERR_INVALID_ARGUMENT;
  }
  if (n > 0 && (center_vox == nullptr || cov_inv == nullptr || rho_r == nullptr || rho_i == nullptr ||
                half_shape == nullptr))
    return GSTV_ERR_INVALID_ARGUMENT;
  if ((n > 0 || count) && (keep_workspace == nullptr || keep_workspace_bytes < gstv_keep_workspace_bytes(n)))
    return GSTV_ERR_WORKSPACE_TOO_SMALL;
  hipStream_t s = static_cast<hipStream_t>(stream);
  KeepLayout ws;
  keep_layout(keep_workspace, n, &ws);
  if (count) geo.live = ws.live;
  int st = launch_prep(center_vox, cov_inv, rho_r, rho_i, rho_stride, half_shape, n, ws.rec, ws.zt, s, radius_vox,
                       &geo, count ? gauss_offsets : nullptr);
  if (st != GSTV_OK || !count) return st;
  return launch_exclusive_scan(gauss_offsets, (int64_t)n + 1, gauss_offsets, ws.sums, total_pairs, s);
}

int gstv_keep_prologue(const float* positions, const float* scales, const float* rotations, const float* rho_r,
                       const float* rho_i, int32_t rho_stride, int32_t n, int32_t d, int32_t h, int32_t w,
                       int32_t tile, float* center_vox, float* radius_vox, float* cov_inv, float* half_shape,
                       int32_t* gauss_offsets, int64_t* total_pairs, void* keep_workspace,
                       size_t keep_workspace_bytes, void* stream) {
  BinGeom geo{};
  int64_t num_tiles;
  if (n < 0 || rho_stride < 1 || half_shape == nullptr || gauss_offsets == nullptr || total_pairs == nullptr ||
      !tile_geometry(d, h, w, tile, geo.nt, &num_tiles))
    return GSTV_ERR_INVALID_ARGUMENT;
  if (n > 0 && (positions == nullptr || scales == nullptr || rotations == nullptr || rho_r == nullptr ||
                rho_i == nullptr || center_vox == nullptr || radius_vox == nullptr || cov_inv == nullptr))
    return GSTV_ERR_INVALID_ARGUMENT;
  if (reinterpret_cast<uintptr_t>(rotations) % 16 != 0) return GSTV_ERR_INVALID_ARGUMENT;
  if (keep_workspace == nullptr || keep_workspace_bytes < gstv_keep_workspace_bytes(n))
    return GSTV_ERR_WORKSPACE_TOO_SMALL;
  geo.size[0] = d;
  geo.size[1] = h;
  geo.size[2] = w;
  geo.tile = tile;
  geo.tile_shift = tile_shift_of(tile);
  hipStream_t s = static_cast<hipStream_t>(stream);
  KeepLayout ws;
  keep_layout(keep_workspace, n, &ws);
  geo.live = ws.live;
  hipLaunchKernelGGL(k_prologue_prep, dim3(ceil_div_i(n + 1, 256)), dim3(256), 0, s, positions, scales, rotations, n,
                     (float)d, (float)h, (float)w, center_vox, radius_vox, cov_inv, half_shape, rho_r, rho_i,
                     (int)rho_stride, ws.rec, ws.zt, geo, gauss_offsets, ws.gacc);
  GSTV_LAUNCH_CHECK("keep_prologue");
  return launch_exclusive_scan(gauss_offsets, (int64_t)n + 1, gauss_offsets, ws.sums, total_pairs, s);
}

int gstv_keep_clear(void* keep_workspace, size_t keep_workspace_bytes, int32_t n, void* stream) {
  if (n < 0) return GSTV_ERR_INVALID_ARGUMENT;
  if (n == 0) return GSTV_OK;
  if (keep_workspace == nullptr || keep_workspace_bytes < gstv_keep_workspace_bytes(n))
    return GSTV_ERR_WORKSPACE_TOO_SMALL;
  KeepLayout ws;
  keep_layout(keep_workspace, n, &ws);
  launch_clear_acc(ws.gacc, n, static_cast<hipStream_t>(stream));
  GSTV_LAUNCH_CHECK("keep_clear");
  return GSTV_OK;
}

int gstv_keep_bin_fill(const void* keep_workspace, size_t keep_workspace_bytes, const float* center_vox,
                       const float* radius_vox, int32_t n, int32_t d, int32_t h, int32_t w, int32_t tile,
                       const int32_t* gauss_offsets, int64_t total_pairs, int32_t* tile_ranges, int32_t* pair_ids,
                       void* workspace, size_t workspace_bytes, void* stream) {
  if (n < 0) return GSTV_ERR_INVALID_ARGUMENT;
  if (n > 0 && (keep_workspace == nullptr || keep_workspace_bytes < gstv_keep_workspace_bytes(n)))
    return GSTV_ERR_WORKSPACE_TOO_SMALL;
  KeepLayout ws{nullptr, nullptr, nullptr, nullptr, nullptr};
  if (n > 0) keep_layout(const_cast<void*>(keep_workspace), n, &ws);
  return bin_fill_impl(ws.rec, ws.live, center_vox, radius_vox, n, d, h, w, tile, gauss_offsets, total_pairs, nullptr,
                       tile_ranges, pair_ids, workspace, workspace_bytes, stream);
}

int gstv_keep_bin_fill_capped(const void* keep_workspace, size_t keep_workspace_bytes, const float* center_vox,
                              const float* radius_vox, int32_t n, int32_t d, int32_t h, int32_t w, int32_t tile,
                              const int32_t* gauss_offsets, const int64_t* total_pairs, int64_t capacity,
                              int32_t* tile_ranges, int32_t* pair_ids, void* workspace, size_t workspace_bytes,
                              void* stream) {
  if (n < 0 || total_pairs == nullptr || capacity < 0) return GSTV_ERR_INVALID_ARGUMENT;
  if (n == 0) capacity = 0;  // no Gaussian, no pair: empty ranges
  if (n > 0 && (keep_workspace == nullptr || keep_workspace_bytes < gstv_keep_workspace_bytes(n)))
    return GSTV_ERR_WORKSPACE_TOO_SMALL;
  KeepLayout ws{nullptr, nullptr, nullptr, nullptr, nullptr};
  if (n > 0) keep_layout(const_cast<void*>(keep_workspace), n, &ws);
  return bin_fill_impl(ws.rec, ws.live, center_vox, radius_vox, n, d, h, w, tile, gauss_offsets, capacity, total_pairs,
                       tile_ranges, pair_ids, workspace, workspace_bytes, stream);
}

int gstv_forward_keep(const void* keep_workspace, size_t keep_workspace_bytes, const int32_t* tile_ranges,
                      int32_t num_tiles, const int32_t* pair_ids, int32_t n, int32_t tile, int32_t d, int32_t h,
                      int32_t w, float* vol_r, float* vol_i, uint32_t* pair_columns, void* stream) {
  int nt[3];
  int st = check_tile_args(tile_ranges, num_tiles, n, tile, d, h, w, nt);
  if (st != GSTV_OK) return st;
  if (vol_r == nullptr || vol_i == nullptr) return GSTV_ERR_INVALID_ARGUMENT;
  if (n > 0 && tile <= 8 && pair_columns == nullptr) return GSTV_ERR_INVALID_ARGUMENT;
  if (n > 0 && (keep_workspace == nullptr || keep_workspace_bytes < gstv_keep_workspace_bytes(n)))
    return GSTV_ERR_WORKSPACE_TOO_SMALL;
  hipStream_t s = static_cast<hipStream_t>(stream);
  KeepLayout ws;
  keep_layout(const_cast<void*>(keep_workspace), n, &ws);
  launch_forward(ws.rec, ws.zt, tile_ranges, num_tiles, pair_ids, tile, d, h, w, nt, vol_r, vol_i, pair_columns, s);
  GSTV_LAUNCH_CHECK("forward_keep");
  return GSTV_OK;
}

int gstv_backward_keep(void* keep_workspace, size_t keep_workspace_bytes, const float* cov_inv,
                       const float* half_shape, const int32_t* tile_ranges, int32_t num_tiles,
                       const int32_t* pair_ids, int32_t n, const float* grad_r, const float* grad_i, int32_t tile,
                       int32_t d, int32_t h, int32_t w, float* grad_center, float* grad_cov_inv,
                       float* grad_rho_r, float* grad_rho_i, int32_t rho_stride, const uint32_t* pair_columns,
                       void* stream) {
  int nt[3];
  int st = check_tile_args(tile_ranges, num_tiles, n, tile, d, h, w, nt);
  if (st != GSTV_OK) return st;
  if (rho_stride < 1) return GSTV_ERR_INVALID_ARGUMENT;
  if (n == 0) return GSTV_OK;
  if (cov_inv == nullptr || half_shape == nullptr || grad_r == nullptr ||
      grad_i == nullptr || grad_center == nullptr || grad_cov_inv == nullptr || grad_rho_r == nullptr ||
      grad_rho_i == nullptr || (tile <= 8 && pair_columns == nullptr))
    return GSTV_ERR_INVALID_ARGUMENT;
  if (keep_workspace == nullptr || keep_workspace_bytes < gstv_keep_workspace_bytes(n))
    return GSTV_ERR_WORKSPACE_TOO_SMALL;
  hipStream_t s = static_cast<hipStream_t>(stream);
  KeepLayout ws;
  keep_layout(keep_workspace, n, &ws);
  launch_backward(ws.rec, ws.zt, ws.gacc, cov_inv, half_shape, tile_ranges, num_tiles, pair_ids, n, grad_r, grad_i,
                  tile, d, h, w, nt, grad_center, grad_cov_inv, grad_rho_r, grad_rho_i, rho_stride, pair_columns, s);
  GSTV_LAUNCH_CHECK("backward_keep");
  return GSTV_OK;
}

int gstv_backward_keep_params(void* keep_workspace, size_t keep_workspace_bytes, const float* cov_inv,
                              const float* half_shape, const float* scales, const float* rotations,
                              const int32_t* tile_ranges, int32_t num_tiles, const int32_t* pair_ids, int32_t n,
                              const float* grad_r, const float* grad_i, int32_t tile, int32_t d, int32_t h,
                              int32_t w, float* grad_positions, float* grad_scales, float* grad_rotations,
                              float* grad_rho_r, float* grad_rho_i, int32_t rho_stride,
                              const uint32_t* pair_columns, void* stream) {
  int nt[3];
  int st = check_tile_args(tile_ranges, num_tiles, n, tile, d, h, w, nt);
  if (st != GSTV_OK) return st;
  if (rho_stride < 1) return GSTV_ERR_INVALID_ARGUMENT;
  if (n == 0) return GSTV_OK;
  if (cov_inv == nullptr || half_shape == nullptr || scales == nullptr ||
      rotations == nullptr || grad_r == nullptr || grad_i == nullptr || grad_positions == nullptr ||
      grad_scales == nullptr || grad_rotations == nullptr || grad_rho_r == nullptr || grad_rho_i == nullptr ||
      (tile <= 8 && pair_columns == nullptr))
    return GSTV_ERR_INVALID_ARGUMENT;
  if (reinterpret_cast<uintptr_t>(rotations) % 16 != 0 || reinterpret_cast<uintptr_t>(grad_rotations) % 16 != 0)
    return GSTV_ERR_INVALID_ARGUMENT;
  if (keep_workspace == nullptr || keep_workspace_bytes < gstv_keep_workspace_bytes(n))
    return GSTV_ERR_WORKSPACE_TOO_SMALL;
  hipStream_t s = static_cast<hipStream_t>(stream);
  KeepLayout ws;
  keep_layout(keep_workspace, n, &ws);
  launch_backward(ws.rec, ws.zt, ws.gacc, cov_inv, half_shape, tile_ranges, num_tiles, pair_ids, n, grad_r, grad_i,
                  tile, d, h, w, nt, nullptr, nullptr, nullptr, nullptr, rho_stride, pair_columns, s);
  hipLaunchKernelGGL(k_finalize_epilogue, dim3(ceil_div_i(n, 256)), dim3(256), 0, s, (const float*)ws.gacc,
                     cov_inv, half_shape, scales, rotations, n, (float)d, (float)h, (float)w, grad_positions,
                     grad_scales, grad_rotations, grad_rho_r, grad_rho_i, (int)rho_stride);
  GSTV_LAUNCH_CHECK("backward_keep_params");
  return GSTV_OK;
}

int gstv_count_evaluations(const float* center_vox, const float* cov_inv, const float* rho_r, const float* rho_i,
                           const int32_t* tile_ranges, int32_t num_tiles, const int32_t* pair_ids,
                           const float* half_shape, int32_t n, int32_t tile, int32_t d, int32_t h, int32_t w,
                           int64_t* counts, void* workspace, size_t workspace_bytes, void* stream) {
  int nt[3];
  int st = check_tile_args(tile_ranges, num_tiles, n, tile, d, h, w, nt);
  if (st != GSTV_OK) return st;
  if (counts == nullptr || half_shape == nullptr) return GSTV_ERR_INVALID_ARGUMENT;
  if (n > 0 && (center_vox == nullptr || cov_inv == nullptr || rho_r == nullptr || rho_i == nullptr))
    return GSTV_ERR_INVALID_ARGUMENT;
  if (n > 0 && (workspace == nullptr || workspace_bytes < gstv_forward_workspace_bytes(n)))
    return GSTV_ERR_WORKSPACE_TOO_SMALL;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, reinterpret_cast<int32_t*>(counts), (int64_t)8, 0);
  GSTV_LAUNCH_CHECK("count_clear");
  GRec* rec;
  ZRec* zt;
  fwd_layout(workspace, n, &rec, &zt);
  st = launch_prep(center_vox, cov_inv, rho_r, rho_i, 1, half_shape, n, rec, zt, s);
  if (st != GSTV_OK) return st;
  const VoxGeom geo = make_vox_geom(d, h, w, tile, nt, num_tiles);
  hipLaunchKernelGGL(k_count, dim3(tile_grid(geo)), dim3(kTileThreads), 0, s, (const GRec*)rec, (const ZRec*)zt,
                     tile_ranges, pair_ids, geo, reinterpret_cast<unsigned long long*>(counts));
  GSTV_LAUNCH_CHECK("count");
  return GSTV_OK;
}

int gstv_prologue(const float* positions, const float* scales, const float* rotations, int32_t n, int32_t d,
                  int32_t h, int32_t w, float* center_vox, float* radius_vox, float* cov_inv, float* half_shape,
                  void* stream) {
  if (n < 0 || d <= 0 || h <= 0 || w <= 0 || half_shape == nullptr) return GSTV_ERR_INVALID_ARGUMENT;
  if (n > 0 && (positions == nullptr || scales == nullptr || rotations == nullptr || center_vox == nullptr ||
                radius_vox == nullptr || cov_inv == nullptr))
    return GSTV_ERR_INVALID_ARGUMENT;
  if (reinterpret_cast<uintptr_t>(rotations) % 16 != 0) return GSTV_ERR_INVALID_ARGUMENT;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_prologue, dim3(ceil_div_i(std::max(n, 1), 256)), dim3(256), 0, s, positions, scales,
                     rotations, n, (float)d, (float)h, (float)w, center_vox, radius_vox, cov_inv, half_shape);
  GSTV_LAUNCH_CHECK("prologue");
  return GSTV_OK;
}

int gstv_epilogue(const float* scales, const float* rotations, const float* grad_center, const float* grad_cov_inv,
                  int32_t n, int32_t d, int32_t h, int32_t w, float* grad_positions, float* grad_scales,
                  float* grad_rotations, void* stream) {
  if (n < 0 || d <= 0 || h <= 0 || w <= 0) return GSTV_ERR_INVALID_ARGUMENT;
  if (n == 0) return GSTV_OK;
  if (scales == nullptr || rotations == nullptr || grad_center == nullptr || grad_cov_inv == nullptr ||
      grad_positions == nullptr || grad_scales == nullptr || grad_rotations == nullptr)
    return GSTV_ERR_INVALID_ARGUMENT;
  if (reinterpret_cast<uintptr_t>(rotations) % 16 != 0 || reinterpret_cast<uintptr_t>(grad_rotations) % 16 != 0)
    return GSTV_ERR_INVALID_ARGUMENT;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_epilogue, dim3(ceil_div_i(n, 256)), dim3(256), 0, s, scales, rotations, grad_center,
                     grad_cov_inv, n, (float)d, (float)h, (float)w, grad_positions, grad_scales, grad_rotations);
  GSTV_LAUNCH_CHECK("epilogue");
  return GSTV_OK;
}

int gstv_profile_begin(void) {
  g_prof.on = true;
  g_prof.used[0] = g_prof.used[1] = 0;
  return GSTV_OK;
}

int gstv_profile_end(double* forward_ms, int32_t* forward_launches, double* backward_ms,
                     int32_t* backward_launches) {
  double total[2] = {0.0, 0.0};
  int status = GSTV_OK;
  for (int kind = 0; kind < 2; ++kind) {
    for (int i = 0; i < g_prof.used[kind]; ++i) {
      float ms = 0.0f;
      int st = hip_status(hipEventSynchronize(g_prof.stop[kind][i]), "profile_sync");
      if (st == GSTV_OK) st = hip_status(hipEventElapsedTime(&ms, g_prof.start[kind][i], g_prof.stop[kind][i]),
                                         "profile_elapsed");
      if (st != GSTV_OK) status = st;
      total[kind] += ms;
    }
  }
  if (forward_ms) *forward_ms = total[0];
  if (forward_launches) *forward_launches = g_prof.used[0];
  if (backward_ms) *backward_ms = total[1];
  if (backward_launches) *backward_launches = g_prof.used[1];
  g_prof.on = false;
  g_prof.used[0] = g_prof.used[1] = 0;
  return status;
}

}  // extern "C"
