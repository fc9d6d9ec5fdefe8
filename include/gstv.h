/*
 * gstv.h - C ABI of the MI355X (gfx950) tile voxelizer.
 *
 * The library renders N anisotropic complex 3D Gaussians into a planar complex
 * volume through 3D voxel tiles, and computes the adjoint (backward) of that
 * rendering.  It replaces the pybind11 extension and the torch-side binning of
 * 3DGSMR's tile voxelizer:
 *
 *   gaussian/tile_voxelizer.py:87-162          build_tile_pairs (cull, expand, sort, ranges)
 *   gaussian/cuda_ext/tile_voxelize_cuda.cpp:33-36   module forward / backward
 *   gaussian/cuda_ext/tile_voxelize_cuda_kernel.cu:262-349  host wrappers
 *
 * Conventions shared by every entry point
 *   - All pointers are device pointers (hipMalloc / torch CUDA tensors), except
 *     where a comment says "host".
 *   - The caller allocates every buffer, including workspaces; the library never
 *     allocates device memory.  Query workspace sizes with the *_workspace_bytes
 *     functions; pass at least that many bytes, aligned to 256 bytes.
 *   - Work is enqueued on `stream` (a hipStream_t, may be NULL for the null
 *     stream) and returns without synchronising.  The device is the caller's
 *     current HIP device.
 *   - The return value is GSTV_OK (0) or a status code; gstv_status_string()
 *     gives the message.  Kernel launch errors are reported as GSTV_ERR_HIP.
 *   - Volumes are row-major [D][H][W]; flat index = d*H*W + h*W + w.
 *   - Tiles are cubes of `tile` voxels; nT* = ceil(size / tile); tile id =
 *     td*(nTH*nTW) + th*nTW + tw.  `tile` is any positive edge; the tile
 *     kernels render a tile above 16 as sub-tiles of at most 16 per axis, each
 *     reading the tile's whole pair list.
 *   - Gaussian parameters are in voxel space: center_vox [N][3] (axis order
 *     D, H, W), cov_inv [N][3][3] row-major (the inverse covariance in units
 *     normalised by half_shape), density planes rho_r [N], rho_i [N].
 *   - half_shape is a device array of 3 floats (D/2, H/2, W/2 for the reference
 *     caller); the kernels read it on the device as the reference kernel does.
 *   - A Gaussian with a non-finite centre coordinate or cov_inv fails the
 *     reference's cutoff test (mahal <= 9, tile_voxelize_cuda_kernel.cu:130,233)
 *     at every voxel: it adds nothing and its gradients are exactly zero.  A
 *     non-finite density, or upstream gradient, reaches exactly the voxels
 *     within its Gaussian's cutoff, as in the reference.
 */
#ifndef GSTV_H
#define GSTV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GSTV_OK 0
#define GSTV_ERR_INVALID_ARGUMENT 1
#define GSTV_ERR_WORKSPACE_TOO_SMALL 2
#define GSTV_ERR_TOO_MANY_PAIRS 3
#define GSTV_ERR_HIP 4
#define GSTV_ERR_TOO_MANY_GAUSSIANS 5

/* Human-readable message for a status code (static storage, never NULL). */
const char* gstv_status_string(int status);

/* Library ABI version (major * 100 + minor). */
int gstv_version(void);

/* ---------------------------------------------------------------- binning --
 * Two-phase binning, as in gaussian/tile_voxelizer.py:87-162: phase one counts
 * the tiles of every Gaussian's 3-sigma box and scans the counts; the caller
 * reads the total (one device-to-host copy, the same sync the reference makes
 * at gaussian/tile_voxelizer.py:118), allocates pair_ids, and calls phase two.
 *
 * Box rule (bit-exact with gaussian/tile_voxelizer.py:97-117): per axis
 *   lo = int(floor(c - r)), hi = int(ceil(c + r)), clamped to [0, size-1];
 *   tile lo/hi = lo / tile, hi / tile, clamped to [0, nT-1].
 * float->int conversion truncates with saturation and maps NaN to 0.
 */

/* Workspace for gstv_bin_count. */
size_t gstv_bin_count_workspace_bytes(int32_t n);

/* Phase one.  Replaces gaussian/tile_voxelizer.py:97-118.
 *   center_vox [n][3], radius_vox [n]           fp32 inputs
 *   gauss_offsets [n+1] int32 out: exclusive scan of per-Gaussian tile counts,
 *                       gauss_offsets[n] = total (valid only if total < 2^31)
 *   total_pairs [1] int64 out (device): the total number of (tile, Gaussian) pairs.
 */
int gstv_bin_count(const float* center_vox, const float* radius_vox, int32_t n,
                   int32_t d, int32_t h, int32_t w, int32_t tile,
                   int32_t* gauss_offsets, int64_t* total_pairs,
                   void* workspace, size_t workspace_bytes, void* stream);

/* Workspace for gstv_bin_fill. */
size_t gstv_bin_fill_workspace_bytes(int32_t n, int64_t total_pairs, int32_t num_tiles);

/* Phase two.  Replaces gaussian/tile_voxelizer.py:120-162.
 *   total_pairs   host value read back from phase one; must be < 2^31
 *                 (GSTV_ERR_TOO_MANY_PAIRS otherwise, as pair offsets are int32
 *                 in the reference, gaussian/tile_voxelizer.py:160-161).
 *   tile_ranges [num_tiles][2] int32 out: [start, end) into pair_ids, (0, 0) for
 *                 tiles without pairs.
 *   pair_ids [total_pairs] int32 out: Gaussian ids sorted by tile id; within a
 *                 tile in ascending Gaussian id (a stable sort of the g-major
 *                 pair list, the canonical order of the reference's argsort).
 */
int gstv_bin_fill(const float* center_vox, const float* radius_vox, int32_t n,
                  int32_t d, int32_t h, int32_t w, int32_t tile,
                  const int32_t* gauss_offsets, int64_t total_pairs,
                  int32_t* tile_ranges, int32_t* pair_ids,
                  void* workspace, size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------- forward --
 * Replaces tile_voxelize_forward_cuda (gaussian/cuda_ext/tile_voxelize_cuda_kernel.cu:262-300)
 * and its kernel (:48-143):
 *   vol[x] = sum over g in tile(x) of rho_g * exp(-q/2) * [q <= 9],
 *   q = d^T M_g d,  d = (x - c_g) / half_shape.
 * Every voxel of vol_r / vol_i [d*h*w] is written (zero where no Gaussian
 * contributes), so the outputs need no prior clearing.  Deterministic.
 *   n           number of Gaussians (rows of center_vox / cov_inv / rho)
 *   num_tiles   rows of tile_ranges; must equal nTD*nTH*nTW for (d, h, w, tile)
 *   pair_ids    read only through tile_ranges: may be null when no tile has
 *               pairs (the data pointer of an empty array), here and in every
 *               other entry point that takes tile lists
 */
size_t gstv_forward_workspace_bytes(int32_t n);

int gstv_forward(const float* center_vox, const float* cov_inv,
                 const float* rho_r, const float* rho_i,
                 const int32_t* tile_ranges, int32_t num_tiles, const int32_t* pair_ids,
                 const float* half_shape, int32_t n,
                 int32_t tile, int32_t d, int32_t h, int32_t w,
                 float* vol_r, float* vol_i,
                 void* workspace, size_t workspace_bytes, void* stream);

/* --------------------------------------------------------------- backward --
 * Replaces tile_voxelize_backward_cuda (gaussian/cuda_ext/tile_voxelize_cuda_kernel.cu:302-349)
 * and its kernel (:145-258).  Given the upstream gradient planes grad_r,
 * grad_i [d*h*w], writes (overwriting, no prior clearing needed):
 *   grad_center [n][3], grad_cov_inv [n][3][3], grad_rho_r [n], grad_rho_i [n]
 * with the reference's formulas: w = exp(-q/2), s = -w/2 (rho_r g_r + rho_i g_i),
 *   grad_rho += w g;  grad_center_i += s * (-2 (M d)_i / hs_i);  grad_M_ij += s d_i d_j.
 * Exponents are recomputed, not stored.  Summation order across tiles is not
 * fixed (fp32 atomics), as in the reference.
 */
size_t gstv_backward_workspace_bytes(int32_t n);

int gstv_backward(const float* center_vox, const float* cov_inv,
                  const float* rho_r, const float* rho_i,
                  const int32_t* tile_ranges, int32_t num_tiles, const int32_t* pair_ids,
                  const float* half_shape, int32_t n,
                  const float* grad_r, const float* grad_i,
                  int32_t tile, int32_t d, int32_t h, int32_t w,
                  float* grad_center, float* grad_cov_inv,
                  float* grad_rho_r, float* grad_rho_i,
                  void* workspace, size_t workspace_bytes, void* stream);

/* ------------------------------------------------- adjoint, no allocation --
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory or hand in a workspace: code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  gstv_backward with the same
 * arguments and results (up to the order of fp32 atomic additions), without
 * the workspace: it keeps its workspace in a static device array of the
 * library, 43.1 MB (GSTV_NOALLOC_MAX_GAUSSIANS x 164 B + 64 KiB) on every
 * GPU it runs on, held until the process ends.
 * It allocates nothing and makes no memset (its kernels rewrite every part of
 * the array they read).
 *   n  at most GSTV_NOALLOC_MAX_GAUSSIANS; larger n returns
 *      GSTV_ERR_TOO_MANY_GAUSSIANS (use gstv_backward).
 * Every call on one GPU shares that GPU's array: calls that overlap on a GPU
 * (concurrent streams or threads) race and give wrong results; calls ordered
 * on one stream are safe. */
#define GSTV_NOALLOC_MAX_GAUSSIANS 262144

int gstv_backward_noalloc(const float* center_vox, const float* cov_inv,
                          const float* rho_r, const float* rho_i,
                          const int32_t* tile_ranges, int32_t num_tiles, const int32_t* pair_ids,
                          const float* half_shape, int32_t n,
                          const float* grad_r, const float* grad_i,
                          int32_t tile, int32_t d, int32_t h, int32_t w,
                          float* grad_center, float* grad_cov_inv,
                          float* grad_rho_r, float* grad_rho_i, void* stream);

/* ------------------------------------------- forward and adjoint of one call --
 * No counterpart in the reference: a fast path for callers that run the
 * adjoint of the same call after the forward (the fused autograd call).  The
 * calls share one caller-allocated keep workspace
 * (gstv_keep_workspace_bytes(n) bytes; keep it, unchanged, between them):
 *   gstv_keep_prepare   builds the Gaussian records that gstv_forward and
 *                       gstv_backward build internally and, when
 *                       gauss_offsets / total_pairs are given, does the work
 *                       of gstv_bin_count restricted to the pairs whose tile
 *                       meets the Gaussian's extent box (the box of its q <= 9
 *                       ellipsoid in the record, widened so that no voxel with
 *                       fp32 q <= 9 lies outside) and, for tile boxes of at
 *                       most 32 tiles (4 per axis), can hold a voxel of that
 *                       widened ellipsoid by its (D, H), (D, W) and (H, W)
 *                       projections over the tile.  Every dropped pair adds
 *                       nothing to the reference's result, so the forward and
 *                       the adjoint on such lists make the same active
 *                       evaluations; the forward's fp32 sums differ in order.
 *   gstv_keep_prologue  gstv_prologue and gstv_keep_prepare (with the counts)
 *                       in one pass; the same outputs.  It also zeroes the
 *                       adjoint's moment accumulator (gstv_keep_clear).
 *   gstv_keep_bin_fill  gstv_bin_fill for those counts (same order within each
 *                       tile: ascending Gaussian id).
 *   gstv_keep_bin_fill_capped
 *                       gstv_keep_bin_fill without the host knowing the total:
 *                       the kernels read it from total_pairs (device memory,
 *                       as gstv_keep_prepare / gstv_keep_prologue wrote it),
 *                       and `capacity` sizes pair_ids (capacity entries) and
 *                       the workspace (gstv_bin_fill_workspace_bytes(n,
 *                       capacity, num_tiles)).  The result equals
 *                       gstv_keep_bin_fill's when total <= capacity; a larger
 *                       total gives empty tile lists (every tile (0, 0)), and
 *                       the caller, who reads the total back once the work
 *                       behind it is queued, redoes the call with an exact
 *                       size.  Replaces the blocking read of the total at
 *                       gaussian/tile_voxelizer.py:118.
 *   gstv_keep_clear     zeroes the adjoint's moment accumulator.  Neither it
 *                       nor the records need binning, so a caller can queue
 *                       them before reading the pair total back to the host.
 *   gstv_forward_keep   gstv_forward on the records; also writes
 *                       pair_columns [total pairs] (uint32, parallel to
 *                       pair_ids): for every pair whose extent box touches its
 *                       tile, the D rows that its column bound leaves in each
 *                       4x4 (H, W) column of the tile (7 bits per column; 0
 *                       for the other pairs).  Required, and written,
 *                       when tile <= 8; ignored for larger tiles.  Any tile
 *                       lists may be given (those of gstv_bin_fill too).
 *   gstv_backward_keep  gstv_backward on the records, the zeroed accumulator
 *                       and pair_columns, for exactly the inputs and lists of
 *                       that forward.  It consumes the accumulator: a second
 *                       adjoint of the same forward needs gstv_keep_clear again.
 *   gstv_backward_keep_params
 *                       gstv_backward_keep followed by gstv_epilogue in one
 *                       pass: gradients of positions, linear scales and
 *                       rotations (as gstv_epilogue; rotations 16-byte
 *                       aligned) instead of grad_center / grad_cov_inv, and
 *                       the density gradient.
 * On the same tile lists, results equal gstv_forward's bit for bit and
 * gstv_backward's up to the order of fp32 atomic additions.
 * rho_stride (>= 1) is the distance, in floats, between consecutive Gaussians'
 * values in rho_r / rho_i and in grad_rho_r / grad_rho_i: 1 for planar arrays,
 * 2 with rho_i = rho_r + 1 (and grad_rho_i = grad_rho_r + 1) for an
 * interleaved complex64 density and its gradient. */
size_t gstv_keep_workspace_bytes(int32_t n);

int gstv_keep_prepare(const float* center_vox, const float* radius_vox, const float* cov_inv,
                      const float* rho_r, const float* rho_i, int32_t rho_stride,
                      const float* half_shape, int32_t n,
                      int32_t d, int32_t h, int32_t w, int32_t tile,
                      int32_t* gauss_offsets, int64_t* total_pairs,
                      void* keep_workspace, size_t keep_workspace_bytes, void* stream);

int gstv_keep_prologue(const float* positions, const float* scales, const float* rotations,
                       const float* rho_r, const float* rho_i, int32_t rho_stride, int32_t n,
                       int32_t d, int32_t h, int32_t w, int32_t tile,
                       float* center_vox, float* radius_vox, float* cov_inv, float* half_shape,
                       int32_t* gauss_offsets, int64_t* total_pairs,
                       void* keep_workspace, size_t keep_workspace_bytes, void* stream);

int gstv_keep_clear(void* keep_workspace, size_t keep_workspace_bytes, int32_t n, void* stream);

int gstv_keep_bin_fill(const void* keep_workspace, size_t keep_workspace_bytes,
                       const float* center_vox, const float* radius_vox, int32_t n,
                       int32_t d, int32_t h, int32_t w, int32_t tile,
                       const int32_t* gauss_offsets, int64_t total_pairs,
                       int32_t* tile_ranges, int32_t* pair_ids,
                       void* workspace, size_t workspace_bytes, void* stream);

int gstv_keep_bin_fill_capped(const void* keep_workspace, size_t keep_workspace_bytes,
                              const float* center_vox, const float* radius_vox, int32_t n,
                              int32_t d, int32_t h, int32_t w, int32_t tile,
                              const int32_t* gauss_offsets, const int64_t* total_pairs, int64_t capacity,
                              int32_t* tile_ranges, int32_t* pair_ids,
                              void* workspace, size_t workspace_bytes, void* stream);

int gstv_forward_keep(const void* keep_workspace, size_t keep_workspace_bytes,
                      const int32_t* tile_ranges, int32_t num_tiles, const int32_t* pair_ids,
                      int32_t n, int32_t tile, int32_t d, int32_t h, int32_t w,
                      float* vol_r, float* vol_i, uint32_t* pair_columns, void* stream);

int gstv_backward_keep(void* keep_workspace, size_t keep_workspace_bytes,
                       const float* cov_inv, const float* half_shape,
                       const int32_t* tile_ranges, int32_t num_tiles, const int32_t* pair_ids,
                       int32_t n, const float* grad_r, const float* grad_i,
                       int32_t tile, int32_t d, int32_t h, int32_t w,
                       float* grad_center, float* grad_cov_inv,
                       float* grad_rho_r, float* grad_rho_i, int32_t rho_stride,
                       const uint32_t* pair_columns, void* stream);

int gstv_backward_keep_params(void* keep_workspace, size_t keep_workspace_bytes,
                              const float* cov_inv, const float* half_shape,
                              const float* scales, const float* rotations,
                              const int32_t* tile_ranges, int32_t num_tiles, const int32_t* pair_ids,
                              int32_t n, const float* grad_r, const float* grad_i,
                              int32_t tile, int32_t d, int32_t h, int32_t w,
                              float* grad_positions, float* grad_scales, float* grad_rotations,
                              float* grad_rho_r, float* grad_rho_i, int32_t rho_stride,
                              const uint32_t* pair_columns, void* stream);

/* ------------------------------------------------------ prologue / epilogue --
 * Replaces the torch prologue of TileVoxelizer.forward (gaussian/tile_voxelizer.py:281-290,
 * with _quaternion_to_rotation_matrix :77-84) and torch autograd through it.
 *   positions [n][3] in [-1, 1], scales [n][3] (linear, the caller's exp(log-scale)),
 *   rotations [n][4] (w, x, y, z; 16-byte aligned).
 * gstv_prologue writes center_vox [n][3] and radius_vox [n] bit-exact with the
 * reference's fp32 operation order, cov_inv [n][3][3] = T T^T with
 * T = R(q / (|q| + 1e-8)) * (1 / (s + 1e-8)) in fp32, and half_shape [3].
 * gstv_epilogue maps grad_center [n][3] and grad_cov_inv [n][3][3] to
 * grad_positions [n][3], grad_scales [n][3] and grad_rotations [n][4]
 * (16-byte aligned); the radius carries no gradient, as in the reference. */
int gstv_prologue(const float* positions, const float* scales, const float* rotations, int32_t n,
                  int32_t d, int32_t h, int32_t w,
                  float* center_vox, float* radius_vox, float* cov_inv, float* half_shape,
                  void* stream);

int gstv_epilogue(const float* scales, const float* rotations,
                  const float* grad_center, const float* grad_cov_inv, int32_t n,
                  int32_t d, int32_t h, int32_t w,
                  float* grad_positions, float* grad_scales, float* grad_rotations,
                  void* stream);

/* ------------------------------------------------------------ measurement --
 * No counterpart in the reference, which has no profiling hooks on this
 * path; these serve the benchmark's roofline accounting.
 */

/* Exact evaluation counts of the tile kernels for the given binning:
 *   counts[0] = E_total   = sum over tiles of pairs(t) * voxels(t)
 *   counts[1] = E_active  = evaluations with q <= 9 (the kernel's own test)
 *               among those the forward kernel keeps after culling
 *   counts[2] = E_culled  = (Gaussian, voxel) evaluations the forward kernel
 *               performs after culling 4x4x4 sub-bricks by the Gaussian's
 *               extent box and its column bound (padding slots excluded)
 *   counts[3] = E_adjoint = (Gaussian, voxel) evaluations the backward kernel
 *               performs: over each pair's live 4x4 (H, W) columns, the D rows
 *               their column bound leaves (padding rows and slots excluded)
 * counts is an int64 [4] device array (overwritten).  Arguments otherwise as
 * for gstv_forward (same workspace size). */
int gstv_count_evaluations(const float* center_vox, const float* cov_inv,
                           const float* rho_r, const float* rho_i,
                           const int32_t* tile_ranges, int32_t num_tiles, const int32_t* pair_ids,
                           const float* half_shape, int32_t n,
                           int32_t tile, int32_t d, int32_t h, int32_t w,
                           int64_t* counts, void* workspace, size_t workspace_bytes, void* stream);

/* Kernel timing with HIP events.  After gstv_profile_begin(), every launch of
 * the forward and backward tile kernels is bracketed by events recorded on the
 * launch stream (at most 1024 launches each).  gstv_profile_end() waits for
 * the recorded events, returns the summed milliseconds and launch counts of
 * each kernel (any pointer may be NULL), and stops recording.  Host-side
 * state is process-wide and not thread-safe; use from one thread. */
int gstv_profile_begin(void);
int gstv_profile_end(double* forward_ms, int32_t* forward_launches,
                     double* backward_ms, int32_t* backward_launches);

#ifdef __cplusplus
}
#endif

#endif /* GSTV_H */
