// mlp.hip - BF16 MFMA GEMMs (FP32 accumulate) for the learned Gaussian initialiser.
//
// One kernel form covers the forward and both backward products of a dense
// layer: C[M,N] = A[M,K] * B[N,K]^T (+ bias[N]) with both operands K-contiguous,
// which is the natural operand form of the gfx950 v_mfma_f32_16x16x32_bf16
// instruction (each lane feeds 8 consecutive k of one row).  The caller forms
// dX = dY * W by passing W^T and dW = dY^T * X by passing transposed copies
// (tiled LDS transposes below).  Block tile 128x128, k-step 32, 4 waves of
// 64x64 (4x4 MFMA tiles each), operands staged through padded LDS rows.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>

#include "gstv_mlp.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kBM = 128, kBN = 128, kBK = 32;
constexpr int kThreads = 256;
constexpr int kPad = 8;  // bf16 elements of padding per LDS row (16 B)

thread_local char g_detail[256] = "";

int hip_status(hipError_t e, const char* where) {
  if (e == hipSuccess) return GSTV_MLP_OK;
  std::snprintf(g_detail, sizeof(g_detail), "%s: %s", where, hipGetErrorString(e));
  return GSTV_MLP_ERR_HIP;
}

// fp32 -> bf16 with round-to-nearest-even; NaN stays NaN.
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x0040u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float apply_act(float v, int act) {
  if (act == GSTV_MLP_ACT_RELU) return fmaxf(v, 0.0f);
  if (act == GSTV_MLP_ACT_GELU) {
    const float c = 0.7978845608028654f;  // sqrt(2/pi), tanh form of GELU
    return 0.5f * v * (1.0f + tanhf(c * (v + 0.044715f * v * v * v)));
  }
  return v;
}

// Loads a 128 x 32 bf16 tile (rows r0.., k0..) of a K-contiguous matrix into
// LDS; out-of-range elements are zero.
__device__ __forceinline__ void load_tile(const uint16_t* __restrict__ src, int64_t ld, int rows, int K, int r0,
                                          int k0, uint16_t (*dst)[kBK + kPad], bool vec_ok) {
  for (int c = threadIdx.x; c < kBM * (kBK / 8); c += kThreads) {
    const int r = c >> 2;
    const int kc = (c & 3) * 8;
    const int gr = r0 + r, gk = k0 + kc;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (gr < rows) {
      const uint16_t* p = src + (int64_t)gr * ld + gk;
      if (vec_ok && gk + 8 <= K) {
        v = *reinterpret_cast<const uint4*>(p);
      } else {
        uint16_t tmp[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) tmp[e] = (gk + e < K) ? p[e] : (uint16_t)0;
        v = make_uint4(tmp[0] | (tmp[1] << 16), tmp[2] | (tmp[3] << 16), tmp[4] | (tmp[5] << 16),
                       tmp[6] | (tmp[7] << 16));
      }
    }
    *reinterpret_cast<uint4*>(&dst[r][kc]) = v;
  }
}

__global__ __launch_bounds__(kThreads) void k_gemm_nt(const uint16_t* __restrict__ A, int64_t lda,
                                                      const uint16_t* __restrict__ B, int64_t ldb,
                                                      const float* __restrict__ bias, int act, int M, int N, int K,
                                                      float* __restrict__ C, int64_t ldc, uint16_t* __restrict__ Cb,
                                                      int64_t ldcb, bool vec_a, bool vec_b) {
  __shared__ __align__(16) uint16_t sA[kBM][kBK + kPad];
  __shared__ __align__(16) uint16_t sB[kBN][kBK + kPad];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.y * kBM, n0 = blockIdx.x * kBN;
  floatx4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};

  const int row = lane & 15;      // operand row inside a 16-row MFMA tile
  const int kq = (lane >> 4) * 8;  // this lane's 8 consecutive k inside the 32-wide step
  for (int k0 = 0; k0 < K; k0 += kBK) {
    load_tile(A, lda, M, K, m0, k0, sA, vec_a);
    load_tile(B, ldb, N, K, n0, k0, sB, vec_b);
    __syncthreads();
    bf16x8 a[4], b[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = *reinterpret_cast<const bf16x8*>(&sA[wm * 64 + i * 16 + row][kq]);
#pragma unroll
    for (int j = 0; j < 4; ++j) b[j] = *reinterpret_cast<const bf16x8*>(&sB[wn * 64 + j * 16 + row][kq]);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  // D layout: lane holds rows 4*(lane/16) .. +3 of column lane%16 of each tile.
  const int rq = (lane >> 4) * 4, col = lane & 15;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = n0 + wn * 64 + j * 16 + col;
      if (n >= N) continue;
      const float bn = bias ? bias[n] : 0.0f;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm * 64 + i * 16 + rq + r;
        if (m >= M) continue;
        const float v = apply_act(acc[i][j][r] + bn, act);
        if (C) C[(int64_t)m * ldc + n] = v;
        if (Cb) Cb[(int64_t)m * ldcb + n] = f32_to_bf16(v);
      }
    }
}

// dst[c][r] = src[r][c] for a rows x cols bf16 matrix, 32x32 tiles through LDS.
__global__ __launch_bounds__(256) void k_transpose(const uint16_t* __restrict__ src, int rows, int cols,
                                                   uint16_t* __restrict__ dst) {
  __shared__ uint16_t t[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 threads
  for (int j = ty; j < 32; j += 8) {
    const int r = by + j, c = bx + tx;
    t[j][tx] = (r < rows && c < cols) ? src[(int64_t)r * cols + c] : (uint16_t)0;
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int c = bx + j, r = by + tx;  // dst row c, column r
    if (c < cols && r < rows) dst[(int64_t)c * rows + r] = t[tx][j];
  }
}

__global__ void k_cast(const float* __restrict__ src, int64_t n, uint16_t* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = f32_to_bf16(src[i]);
}

}  // namespace

extern "C" {

const char* gstv_mlp_status_string(int status) {
  switch (status) {
    case GSTV_MLP_OK: return "success";
    case GSTV_MLP_ERR_INVALID_ARGUMENT: return "invalid argument";
    case GSTV_MLP_ERR_HIP: return g_detail[0] ? g_detail : "HIP runtime error";
    default: return "unknown status";
  }
}

int gstv_mlp_gemm_nt_bf16(const void* a, int64_t lda, const void* b, int64_t ldb, const float* bias, int act, int m,
                          int n, int k, float* c, int64_t ldc, void* c_bf16, int64_t ldcb, void* stream) {
  if (m < 0 || n < 0 || k < 0 || (m > 0 && n > 0 && k > 0 && (a == nullptr || b == nullptr)) ||
      (c == nullptr && c_bf16 == nullptr) || lda < k || ldb < k || (c && ldc < n) || (c_bf16 && ldcb < n) ||
      act < 0 || act > 2)
    return GSTV_MLP_ERR_INVALID_ARGUMENT;
  if (m == 0 || n == 0) return GSTV_MLP_OK;
  const bool vec_a = (reinterpret_cast<uintptr_t>(a) % 16 == 0) && (lda % 8 == 0);
  const bool vec_b = (reinterpret_cast<uintptr_t>(b) % 16 == 0) && (ldb % 8 == 0);
  dim3 grid((n + kBN - 1) / kBN, (m + kBM - 1) / kBM);
  hipLaunchKernelGGL(k_gemm_nt, grid, dim3(kThreads), 0, static_cast<hipStream_t>(stream),
                     static_cast<const uint16_t*>(a), lda, static_cast<const uint16_t*>(b), ldb, bias, act, m, n, k,
                     c, ldc, static_cast<uint16_t*>(c_bf16), ldcb, vec_a, vec_b);
  return hip_status(hipGetLastError(), "gemm_nt_bf16");
}

int gstv_mlp_transpose_bf16(const void* src, int rows, int cols, void* dst, void* stream) {
  if (rows < 0 || cols < 0 || ((rows > 0 && cols > 0) && (src == nullptr || dst == nullptr)))
    return GSTV_MLP_ERR_INVALID_ARGUMENT;
  if (rows == 0 || cols == 0) return GSTV_MLP_OK;
  dim3 grid((cols + 31) / 32, (rows + 31) / 32);
  hipLaunchKernelGGL(k_transpose, grid, dim3(256), 0, static_cast<hipStream_t>(stream),
                     static_cast<const uint16_t*>(src), rows, cols, static_cast<uint16_t*>(dst));
  return hip_status(hipGetLastError(), "transpose_bf16");
}

int gstv_mlp_cast_bf16(const float* src, int64_t n, void* dst, void* stream) {
  if (n < 0 || (n > 0 && (src == nullptr || dst == nullptr))) return GSTV_MLP_ERR_INVALID_ARGUMENT;
  if (n == 0) return GSTV_MLP_OK;
  hipLaunchKernelGGL(k_cast, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, static_cast<hipStream_t>(stream), src, n,
                     static_cast<uint16_t*>(dst));
  return hip_status(hipGetLastError(), "cast_bf16");
}

}  // extern "C"
