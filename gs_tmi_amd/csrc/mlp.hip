// mlp.hip - BF16 MFMA GEMMs (FP32 accumulate) for the learned Gaussian initialiser.
//
// One kernel form covers the forward and both backward products of a dense
// layer: C[M,N] = A[M,K] * B[N,K]^T (+ bias[N]) with both operands K-contiguous,
// which is the natural operand form of the gfx950 v_mfma_f32_16x16x32_bf16
// instruction (each lane feeds 8 consecutive k of one row).  The caller forms
// dX = dY * W by passing W^T and dW = dY^T * X by passing transposed copies
// (tiled LDS transposes below).  Block tiles 128x128 or 64x64 (chosen so the
// grid fills the chip), 64-wide k stages double-buffered in padded LDS rows,
// 4 waves each owning a quarter of the block tile.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>

#include "gstv_mlp.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kBK = 64;                 // k per LDS stage
constexpr int kPad = 8;                 // bf16 elements of padding per LDS row (16 B): conflict-free b128 reads
constexpr int kLdk = kBK + kPad;

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

// Global -> registers: this thread's 16-byte chunks of a ROWS x 64 tile of a
// K-contiguous bf16 matrix (chunk c: row c / 8, k 8 * (c % 8)); out-of-range
// elements are zero.
template <int ROWS, int NT>
__device__ __forceinline__ void load_chunks(const uint16_t* __restrict__ src, int64_t ld, int rows, int K, int r0,
                                            int k0, bool vec_ok, uint4 (&reg)[ROWS * kBK / 8 / NT]) {
#pragma unroll
  for (int i = 0; i < ROWS * kBK / 8 / NT; ++i) {
    const int c = threadIdx.x + i * NT;
    const int gr = r0 + (c >> 3), gk = k0 + (c & 7) * 8;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (gr < rows) {
      const uint16_t* p = src + (int64_t)gr * ld + gk;
      if (vec_ok && gk + 8 <= K) {
        v = *reinterpret_cast<const uint4*>(p);
      } else {
        uint16_t t[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) t[e] = (gk + e < K) ? p[e] : (uint16_t)0;
        v = make_uint4(t[0] | (t[1] << 16), t[2] | (t[3] << 16), t[4] | (t[5] << 16), t[6] | (t[7] << 16));
      }
    }
    reg[i] = v;
  }
}

template <int ROWS, int NT>
__device__ __forceinline__ void store_chunks(uint16_t (*dst)[kLdk], const uint4 (&reg)[ROWS * kBK / 8 / NT]) {
#pragma unroll
  for (int i = 0; i < ROWS * kBK / 8 / NT; ++i) {
    const int c = threadIdx.x + i * NT;
    *reinterpret_cast<uint4*>(&dst[c >> 3][(c & 7) * 8]) = reg[i];
  }
}

// C[M,N] = act(A[M,K] B[N,K]^T + bias).  Block tile BM x BN, GM x GN waves of
// (BM/GM) x (BN/GN), each a grid of 16 x 16 MFMA tiles.  The k loop runs over
// 64-wide stages held in two LDS buffers: the next stage's global loads are
// issued into registers before the current stage's MFMAs and stored to the
// other buffer after them, so one barrier per stage suffices.
template <int BM, int BN, int GM = 2, int GN = 2>
__global__ __launch_bounds__(GM * GN * 64) void k_gemm_nt(const uint16_t* __restrict__ A, int64_t lda,
                                                      const uint16_t* __restrict__ B, int64_t ldb,
                                                      const float* __restrict__ bias, int act, int M, int N, int K,
                                                      float* __restrict__ C, int64_t ldc, uint16_t* __restrict__ Cb,
                                                      int64_t ldcb, bool vec_a, bool vec_b) {
  constexpr int NT = GM * GN * 64;
  constexpr int WM = BM / GM, WN = BN / GN, TM = WM / 16, TN = WN / 16;
  __shared__ __align__(16) uint16_t sA[2][BM][kLdk];
  __shared__ __align__(16) uint16_t sB[2][BN][kLdk];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave / GN, wn = wave % GN;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  floatx4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};

  uint4 ra[BM * kBK / 8 / NT], rb[BN * kBK / 8 / NT];
  const int nk = (K + kBK - 1) / kBK;
  load_chunks<BM, NT>(A, lda, M, K, m0, 0, vec_a, ra);
  load_chunks<BN, NT>(B, ldb, N, K, n0, 0, vec_b, rb);
  store_chunks<BM, NT>(sA[0], ra);
  store_chunks<BN, NT>(sB[0], rb);
  __syncthreads();

  const int row = lane & 15;       // operand row inside a 16-row MFMA tile
  const int kq = (lane >> 4) * 8;  // this lane's 8 consecutive k inside a 32-wide MFMA step
  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    const bool more = kt + 1 < nk;
    if (more) {
      load_chunks<BM, NT>(A, lda, M, K, m0, (kt + 1) * kBK, vec_a, ra);
      load_chunks<BN, NT>(B, ldb, N, K, n0, (kt + 1) * kBK, vec_b, rb);
    }
#pragma unroll
    for (int ks = 0; ks < kBK; ks += 32) {
      bf16x8 a[TM], b[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) a[i] = *reinterpret_cast<const bf16x8*>(&sA[buf][wm * WM + i * 16 + row][ks + kq]);
#pragma unroll
      for (int j = 0; j < TN; ++j) b[j] = *reinterpret_cast<const bf16x8*>(&sB[buf][wn * WN + j * 16 + row][ks + kq]);
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    if (more) {
      store_chunks<BM, NT>(sA[buf ^ 1], ra);
      store_chunks<BN, NT>(sB[buf ^ 1], rb);
    }
    __syncthreads();
  }
  // D layout: lane holds rows 4*(lane/16) .. +3 of column lane%16 of each tile.
  const int rq = (lane >> 4) * 4, col = lane & 15;
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int n = n0 + wn * WN + j * 16 + col;
      if (n >= N) continue;
      const float bn = bias ? bias[n] : 0.0f;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm * WM + i * 16 + rq + r;
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
  // The largest block tile that still gives at least one block per CU:
  // 256 x 256 (16 waves of 64 x 64; 1.0 PFLOP/s at 4096^3 against 0.68 for
  // 128 x 128 with 4 waves), 128 x 128 (8 waves of 32 x 64; 474 against 410
  // TFLOP/s with 4 waves at 16384 x 512 x 1024), else 64 x 64 (8 waves of
  // 32 x 16; 158 against 137 TFLOP/s with 4 waves at 2048 x 512 x 1024).
  const int64_t huge_blocks = (int64_t)((n + 255) / 256) * ((m + 255) / 256);
  const int64_t big_blocks = (int64_t)((n + 127) / 128) * ((m + 127) / 128);
  const hipStream_t st = static_cast<hipStream_t>(stream);
  const uint16_t* pa = static_cast<const uint16_t*>(a);
  const uint16_t* pb = static_cast<const uint16_t*>(b);
  uint16_t* pcb = static_cast<uint16_t*>(c_bf16);
  if (huge_blocks >= 256) {
    dim3 grid((n + 255) / 256, (m + 255) / 256);
    hipLaunchKernelGGL((k_gemm_nt<256, 256, 4, 4>), grid, dim3(16 * 64), 0, st, pa, lda, pb, ldb, bias, act, m, n, k,
                       c, ldc, pcb, ldcb, vec_a, vec_b);
  } else if (big_blocks >= 256) {
    dim3 grid((n + 127) / 128, (m + 127) / 128);
    hipLaunchKernelGGL((k_gemm_nt<128, 128, 4, 2>), grid, dim3(8 * 64), 0, st, pa, lda, pb, ldb, bias, act, m, n, k, c,
                       ldc, pcb, ldcb, vec_a, vec_b);
  } else {
    dim3 grid((n + 63) / 64, (m + 63) / 64);
    hipLaunchKernelGGL((k_gemm_nt<64, 64, 2, 4>), grid, dim3(512), 0, st, pa, lda, pb, ldb, bias, act, m, n, k, c,
                       ldc, pcb, ldcb, vec_a, vec_b);
  }
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
