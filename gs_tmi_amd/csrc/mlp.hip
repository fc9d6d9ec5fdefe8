// mlp.hip - BF16 MFMA GEMMs (FP32 accumulate) for the learned Gaussian initialiser.
//
// One kernel form covers the forward and both backward products of a dense
// layer: C[M,N] = A[M,K] * B[N,K]^T (+ bias[N]) with both operands K-contiguous,
// which is the natural operand form of the gfx950 v_mfma_f32_16x16x32_bf16
// instruction (each lane feeds 8 consecutive k of one row).  The caller forms
// dX = dY * W by passing W^T and dW = dY^T * X by passing transposed copies.
// Those copies come out of the kernels that produce the operands anyway: the
// GEMM epilogue can write its bf16 result transposed as well, and the fp32 ->
// bf16 cast can write a transposed copy beside the plain one; the epilogue can
// also apply a ReLU's gradient mask (the backward of a fused bias + ReLU layer).
// Block tiles 256x256, 128x128 or 64x64 (the largest that still fills the
// chip), 64-wide k stages double-buffered in padded LDS rows, each wave owning
// a rectangle of 16x16 MFMA tiles of the block tile.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>

#include "gstv_mlp.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kPad = 8;  // bf16 elements of padding per LDS row (16 B): conflict-free b128 reads

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

// Global -> registers: this thread's 16-byte chunks of a ROWS x BK tile of a
// K-contiguous bf16 matrix (chunk c: row c / (BK/8), k 8 * (c % (BK/8)));
// out-of-range elements are zero.
template <int ROWS, int NT, int BK>
__device__ __forceinline__ void load_chunks(const uint16_t* __restrict__ src, int64_t ld, int rows, int K, int r0,
                                            int k0, bool vec_ok, uint4 (&reg)[ROWS * BK / 8 / NT]) {
  constexpr int kCpr = BK / 8;  // chunks per row
#pragma unroll
  for (int i = 0; i < ROWS * BK / 8 / NT; ++i) {
    const int c = threadIdx.x + i * NT;
    const int gr = r0 + c / kCpr, gk = k0 + (c % kCpr) * 8;
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

template <int ROWS, int NT, int BK>
__device__ __forceinline__ void store_chunks(uint16_t (*dst)[BK + kPad], const uint4 (&reg)[ROWS * BK / 8 / NT]) {
  constexpr int kCpr = BK / 8;
#pragma unroll
  for (int i = 0; i < ROWS * BK / 8 / NT; ++i) {
    const int c = threadIdx.x + i * NT;
    *reinterpret_cast<uint4*>(&dst[c / kCpr][(c % kCpr) * 8]) = reg[i];
  }
}

// Epilogue options beyond bias and activation.
struct Epi {
  const uint16_t* mask;  // bf16 [M][N] (row stride ldm): zero every result whose mask entry is not > 0
  int64_t ldm;
  uint16_t* ct;          // bf16 copy of the result, transposed: ct[n * ldct + m]
  int64_t ldct;
};

// A bf16 bit pattern that is a positive number (not +0, not NaN).
__device__ __forceinline__ bool bf16_positive(uint16_t h) { return (int16_t)h > 0 && h <= 0x7f80u; }

// C[M,N] = act(A[M,K] B[N,K]^T + bias).  Block tile BM x BN, GM x GN waves of
// (BM/GM) x (BN/GN), each a grid of 16 x 16 MFMA tiles.  The k loop runs over
// BK-wide stages held in two LDS buffers: the next stage's global loads are
// issued into registers before the current stage's MFMAs and stored to the
// other buffer after them, so one barrier per stage suffices.
template <int BM, int BN, int GM, int GN, int BK>
__global__ __launch_bounds__(GM * GN * 64) void k_gemm_nt(const uint16_t* __restrict__ A, int64_t lda,
                                                      const uint16_t* __restrict__ B, int64_t ldb,
                                                      const float* __restrict__ bias, int act, int M, int N, int K,
                                                      float* __restrict__ C, int64_t ldc, uint16_t* __restrict__ Cb,
                                                      int64_t ldcb, bool vec_a, bool vec_b, Epi epi) {
  constexpr int NT = GM * GN * 64;
  constexpr int WM = BM / GM, WN = BN / GN, TM = WM / 16, TN = WN / 16;
  static_assert(TM >= 1 && TN >= 1 && BM * BK / 8 % NT == 0 && BN * BK / 8 % NT == 0, "tile shape");
  __shared__ __align__(16) uint16_t sA[2][BM][BK + kPad];
  __shared__ __align__(16) uint16_t sB[2][BN][BK + kPad];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave / GN, wn = wave % GN;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  floatx4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};

  uint4 ra[BM * BK / 8 / NT], rb[BN * BK / 8 / NT];
  const int nk = (K + BK - 1) / BK;
  load_chunks<BM, NT, BK>(A, lda, M, K, m0, 0, vec_a, ra);
  load_chunks<BN, NT, BK>(B, ldb, N, K, n0, 0, vec_b, rb);
  store_chunks<BM, NT, BK>(sA[0], ra);
  store_chunks<BN, NT, BK>(sB[0], rb);
  __syncthreads();

  const int row = lane & 15;       // operand row inside a 16-row MFMA tile
  const int kq = (lane >> 4) * 8;  // this lane's 8 consecutive k inside a 32-wide MFMA step
  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    const bool more = kt + 1 < nk;
    if (more) {
      load_chunks<BM, NT, BK>(A, lda, M, K, m0, (kt + 1) * BK, vec_a, ra);
      load_chunks<BN, NT, BK>(B, ldb, N, K, n0, (kt + 1) * BK, vec_b, rb);
    }
#pragma unroll
    for (int ks = 0; ks < BK; ks += 32) {
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
      store_chunks<BM, NT, BK>(sA[buf ^ 1], ra);
      store_chunks<BN, NT, BK>(sB[buf ^ 1], rb);
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
      const int mb = m0 + wm * WM + i * 16 + rq;
      uint16_t tb[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = mb + r;
        tb[r] = 0;
        if (m >= M) continue;
        float v = apply_act(acc[i][j][r] + bn, act);
        if (epi.mask != nullptr && !bf16_positive(epi.mask[(int64_t)m * epi.ldm + n])) v = 0.0f;
        if (C) C[(int64_t)m * ldc + n] = v;
        tb[r] = f32_to_bf16(v);
        if (Cb) Cb[(int64_t)m * ldcb + n] = tb[r];
      }
      if (epi.ct != nullptr) {
        // The lane's 4 rows are 4 consecutive entries of row n of the transposed copy.
        uint16_t* dst = epi.ct + (int64_t)n * epi.ldct + mb;
        if (mb + 3 < M && (reinterpret_cast<uintptr_t>(dst) & 7) == 0) {
          *reinterpret_cast<uint2*>(dst) =
              make_uint2(tb[0] | ((uint32_t)tb[1] << 16), tb[2] | ((uint32_t)tb[3] << 16));
        } else {
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (mb + r < M) dst[r] = tb[r];
        }
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

// dst = bf16(src) (when dst is given) and dst_t = its transpose, for a
// rows x cols fp32 matrix, in one pass: 32 x 32 tiles through LDS, both
// writes coalesced.
__global__ __launch_bounds__(256) void k_cast_t(const float* __restrict__ src, int rows, int cols,
                                                uint16_t* __restrict__ dst, uint16_t* __restrict__ dst_t) {
  __shared__ uint16_t t[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 threads
  for (int j = ty; j < 32; j += 8) {
    const int r = by + j, c = bx + tx;
    uint16_t h = 0;
    if (r < rows && c < cols) {
      h = f32_to_bf16(src[(int64_t)r * cols + c]);
      if (dst != nullptr) dst[(int64_t)r * cols + c] = h;
    }
    t[j][tx] = h;
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int c = bx + j, r = by + tx;  // dst_t row c, column r
    if (c < cols && r < rows) dst_t[(int64_t)c * rows + r] = t[tx][j];
  }
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

int gstv_mlp_gemm_nt_bf16_ex(const void* a, int64_t lda, const void* b, int64_t ldb, const float* bias, int act,
                             const void* relu_mask, int64_t ldm, int m, int n, int k, float* c, int64_t ldc,
                             void* c_bf16, int64_t ldcb, void* c_bf16_t, int64_t ldct, void* stream) {
  if (m < 0 || n < 0 || k < 0 || (m > 0 && n > 0 && k > 0 && (a == nullptr || b == nullptr)) ||
      (c == nullptr && c_bf16 == nullptr && c_bf16_t == nullptr) || lda < k || ldb < k || (c && ldc < n) ||
      (c_bf16 && ldcb < n) || (c_bf16_t && ldct < m) || (relu_mask && ldm < n) || act < 0 || act > 2)
    return GSTV_MLP_ERR_INVALID_ARGUMENT;
  const Epi epi{static_cast<const uint16_t*>(relu_mask), ldm, static_cast<uint16_t*>(c_bf16_t), ldct};
  if (m == 0 || n == 0) return GSTV_MLP_OK;
  const bool vec_a = (reinterpret_cast<uintptr_t>(a) % 16 == 0) && (lda % 8 == 0);
  const bool vec_b = (reinterpret_cast<uintptr_t>(b) % 16 == 0) && (ldb % 8 == 0);
  // Block tile: 256 x 256 (16 waves of 64 x 64) or 128 x 128 (8 waves of
  // 32 x 64) when that still gives at least one block per CU (1.0 PFLOP/s at
  // 4096^3; 16384 x 512 x 1024 in 35 us).  Smaller products are latency-bound
  // on their chain of k stages, so they take 128-wide stages: 64 x 64 (8 waves
  // of 32 x 16) when that fills the chip (2048 x 512 x 1024: 11.2 us against
  // 13.8 with 64-wide stages), else 32 x 32 (4 waves of 16 x 16; the weight
  // gradients 512 x {512, 1024} x 2048: 10.5 / 12.5 us against 18.3 / 18.6
  // with 64 x 64 tiles).  Kernel times from a sweep of 12 tile shapes at the
  // learned initialiser's products, rocprofv3 kernel trace.
  const int64_t huge_blocks = (int64_t)((n + 255) / 256) * ((m + 255) / 256);
  const int64_t big_blocks = (int64_t)((n + 127) / 128) * ((m + 127) / 128);
  const int64_t mid_blocks = (int64_t)((n + 63) / 64) * ((m + 63) / 64);
  const hipStream_t st = static_cast<hipStream_t>(stream);
  const uint16_t* pa = static_cast<const uint16_t*>(a);
  const uint16_t* pb = static_cast<const uint16_t*>(b);
  uint16_t* pcb = static_cast<uint16_t*>(c_bf16);
  if (huge_blocks >= 256) {
    dim3 grid((n + 255) / 256, (m + 255) / 256);
    hipLaunchKernelGGL((k_gemm_nt<256, 256, 4, 4, 64>), grid, dim3(16 * 64), 0, st, pa, lda, pb, ldb, bias, act, m, n,
                       k, c, ldc, pcb, ldcb, vec_a, vec_b, epi);
  } else if (big_blocks >= 256) {
    dim3 grid((n + 127) / 128, (m + 127) / 128);
    hipLaunchKernelGGL((k_gemm_nt<128, 128, 4, 2, 64>), grid, dim3(8 * 64), 0, st, pa, lda, pb, ldb, bias, act, m, n,
                       k, c, ldc, pcb, ldcb, vec_a, vec_b, epi);
  } else if (mid_blocks >= 256) {
    dim3 grid((n + 63) / 64, (m + 63) / 64);
    hipLaunchKernelGGL((k_gemm_nt<64, 64, 2, 4, 128>), grid, dim3(8 * 64), 0, st, pa, lda, pb, ldb, bias, act, m, n,
                       k, c, ldc, pcb, ldcb, vec_a, vec_b, epi);
  } else {
    dim3 grid((n + 31) / 32, (m + 31) / 32);
    hipLaunchKernelGGL((k_gemm_nt<32, 32, 2, 2, 128>), grid, dim3(4 * 64), 0, st, pa, lda, pb, ldb, bias, act, m, n,
                       k, c, ldc, pcb, ldcb, vec_a, vec_b, epi);
  }
  return hip_status(hipGetLastError(), "gemm_nt_bf16");
}

int gstv_mlp_gemm_nt_bf16(const void* a, int64_t lda, const void* b, int64_t ldb, const float* bias, int act, int m,
                          int n, int k, float* c, int64_t ldc, void* c_bf16, int64_t ldcb, void* stream) {
  if (c == nullptr && c_bf16 == nullptr) return GSTV_MLP_ERR_INVALID_ARGUMENT;
  return gstv_mlp_gemm_nt_bf16_ex(a, lda, b, ldb, bias, act, nullptr, 0, m, n, k, c, ldc, c_bf16, ldcb, nullptr, 0,
                                  stream);
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

int gstv_mlp_cast_t_bf16(const float* src, int rows, int cols, void* dst, void* dst_t, void* stream) {
  if (rows < 0 || cols < 0 || ((rows > 0 && cols > 0) && (src == nullptr || dst_t == nullptr)))
    return GSTV_MLP_ERR_INVALID_ARGUMENT;
  if (rows == 0 || cols == 0) return GSTV_MLP_OK;
  dim3 grid((cols + 31) / 32, (rows + 31) / 32);
  hipLaunchKernelGGL(k_cast_t, grid, dim3(256), 0, static_cast<hipStream_t>(stream), src, rows, cols,
                     static_cast<uint16_t*>(dst), static_cast<uint16_t*>(dst_t));
  return hip_status(hipGetLastError(), "cast_t_bf16");
}

int gstv_mlp_cast_bf16(const float* src, int64_t n, void* dst, void* stream) {
  if (n < 0 || (n > 0 && (src == nullptr || dst == nullptr))) return GSTV_MLP_ERR_INVALID_ARGUMENT;
  if (n == 0) return GSTV_MLP_OK;
  hipLaunchKernelGGL(k_cast, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, static_cast<hipStream_t>(stream), src, n,
                     static_cast<uint16_t*>(dst));
  return hip_status(hipGetLastError(), "cast_bf16");
}

}  // extern "C"
