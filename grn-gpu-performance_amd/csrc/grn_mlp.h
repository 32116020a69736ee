// grn_mlp.h -- the surrogate network's shapes, kernel arguments and the
// host-side launchers of grn_mlp.hip.
#ifndef GRN_MLP_H
#define GRN_MLP_H

#include <hip/hip_runtime.h>

// 468 encoded features (12 genes x 3 parameters + 12 x 12 links x 3), zero-
// padded to the 512-wide layers; four hidden layers of 512 with GELU; one
// scalar output.
#define MLP_FEATURES 468
#define MLP_GENES 12
#define MLP_WIDTH 512
#define MLP_HIDDEN_LAYERS 4
#define MLP_ROW_ALIGN 128  // GEMM tile rows: batches are padded to a multiple

// Flat FP32 parameter vector: W_l^T (out x in) for l = 1..4, then b_1..b_4,
// then the output weights w (512) and bias (1).
#define MLP_MAT (MLP_WIDTH * MLP_WIDTH)
#define MLP_OFF_BIAS ((long)MLP_HIDDEN_LAYERS * MLP_MAT)
#define MLP_OFF_OUT_W (MLP_OFF_BIAS + (long)MLP_HIDDEN_LAYERS * MLP_WIDTH)
#define MLP_OFF_OUT_B (MLP_OFF_OUT_W + MLP_WIDTH)
#define MLP_N_PARAMS (MLP_OFF_OUT_B + 1)

enum { EPI_F32 = 0, EPI_FWD = 1, EPI_BWD = 2 };

// C[m][n] = sum_k A[m][k] Bt[n][k]; M, N multiples of 128, K of 32.
//   EPI_F32: C (fp32) = the product.
//   EPI_FWD: z = product + bias[n]; C = z (optional); H = gelu(z) (bf16),
//            Ht = its transpose (optional).
//   EPI_BWD: d = product * gelu'(Zprev); H = d (bf16), Ht = transpose (optional).
struct MlpGemm {
  const __bf16 *A;
  const __bf16 *Bt;
  int M, N, K;
  float *C;
  const float *bias;
  __bf16 *H;
  __bf16 *Ht;
  const float *Zprev;
  // EPI_F32 only: ksplit > 1 slices K over that many workgroups per tile;
  // each slice writes its partial product to Cpart[slice][M][N] and a second
  // kernel sums the slices into C in slice order (deterministic).  K must be
  // a multiple of 32 * ksplit.
  int ksplit;
  float *Cpart;
};

struct MlpAdam {
  float *p, *m, *v;
  const float *grad;
  long n;
  float beta1, beta2, lr_t, eps_t;  // lr and eps with Adam's bias corrections folded in
  __bf16 *wt;                       // [4][out][in]
  __bf16 *w;                        // [4][in][out]
  __bf16 *w_out;                    // [512]
  long out_w_offset;
};

hipError_t grn_mlp_gemm(int epi, const MlpGemm &g, hipStream_t s);
hipError_t grn_mlp_out_forward(const __bf16 *H, const __bf16 *w, const float *b0, float *y, int rows, hipStream_t s);
// pair_w > 0: rows come in (child, parent) pairs and the loss adds the
// pairwise difference term (rows must then be even).
hipError_t grn_mlp_out_backward(const float *y, const float *t, const float *mask, const __bf16 *w, const float *Z,
                                float inv_n, float pair_w, int rows, float *dy, float *sq, __bf16 *dZ, __bf16 *dZt,
                                hipStream_t s);
hipError_t grn_mlp_row_dot(const __bf16 *X, const float *v, int rows, int cols, float *out, hipStream_t s);
hipError_t grn_mlp_sum(const float *x, int n, float *out, hipStream_t s);
hipError_t grn_mlp_adam(const MlpAdam &a, hipStream_t s);
hipError_t grn_mlp_gather(const __bf16 *feat, const float *target, const int *idx, int rows, float mu,
                          float inv_sigma, __bf16 *X, __bf16 *Xt, float *t, hipStream_t s);

#endif
