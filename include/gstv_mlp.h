/*
 * gstv_mlp.h - C ABI of the BF16 MFMA GEMMs of the learned Gaussian initialiser.
 *
 * No counterpart in the reference: 3DGSMR initialises Gaussians only from the
 * zero-filled image (gaussian/gaussian_model.py:308-355).  The learned
 * initialiser is an MLP over 8^3 patches of the zero-filled image; these
 * entry points are its dense-layer products on gfx950 matrix cores.
 *
 * Conventions: device pointers; bf16 buffers are uint16 bit patterns; the
 * caller allocates everything; work is enqueued on `stream` (hipStream_t).
 */
#ifndef GSTV_MLP_H
#define GSTV_MLP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GSTV_MLP_OK 0
#define GSTV_MLP_ERR_INVALID_ARGUMENT 1
#define GSTV_MLP_ERR_HIP 4

#define GSTV_MLP_ACT_NONE 0
#define GSTV_MLP_ACT_RELU 1
#define GSTV_MLP_ACT_GELU 2

const char* gstv_mlp_status_string(int status);

/* C[m][n] = act( sum_k A[m][k] * B[n][k] + bias[n] ), bf16 operands, fp32
 * accumulation (v_mfma_f32_16x16x32_bf16).  A is m x k with row stride lda,
 * B is n x k with row stride ldb (both K-contiguous); bias may be NULL.
 * Writes the fp32 result to C (row stride ldc) and/or a bf16 copy to C_bf16
 * (row stride ldcb); either output pointer may be NULL, not both. */
int gstv_mlp_gemm_nt_bf16(const void* a, int64_t lda, const void* b, int64_t ldb,
                          const float* bias, int act, int m, int n, int k,
                          float* c, int64_t ldc, void* c_bf16, int64_t ldcb, void* stream);

/* gstv_mlp_gemm_nt_bf16 with two more epilogue outputs / options:
 *   relu_mask  bf16 m x n (row stride ldm) or NULL: every result whose mask
 *              entry is not a positive number is zeroed (after bias and act):
 *              the backward of a ReLU whose output is the mask;
 *   c_bf16_t   bf16 n x m (row stride ldct >= m) or NULL: the bf16 result,
 *              transposed (c_bf16_t[j * ldct + i] = bf16(C[i][j])), the
 *              operand form the next weight-gradient product reads.
 * At least one of c, c_bf16, c_bf16_t must be given. */
int gstv_mlp_gemm_nt_bf16_ex(const void* a, int64_t lda, const void* b, int64_t ldb,
                             const float* bias, int act, const void* relu_mask, int64_t ldm,
                             int m, int n, int k, float* c, int64_t ldc,
                             void* c_bf16, int64_t ldcb, void* c_bf16_t, int64_t ldct, void* stream);

/* dst (rows x cols, may be NULL) = bf16(src) and dst_t (cols x rows) = its
 * transpose, in one pass over the fp32 rows x cols matrix src (RNE). */
int gstv_mlp_cast_t_bf16(const float* src, int rows, int cols, void* dst, void* dst_t, void* stream);

/* dst (cols x rows) = transpose of src (rows x cols), bf16. */
int gstv_mlp_transpose_bf16(const void* src, int rows, int cols, void* dst, void* stream);

/* dst[i] = bf16(src[i]) with round-to-nearest-even. */
int gstv_mlp_cast_bf16(const float* src, int64_t n, void* dst, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GSTV_MLP_H */
