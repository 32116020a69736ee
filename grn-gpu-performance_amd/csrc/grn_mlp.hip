// grn_mlp.hip -- kernels of the surrogate network that pre-screens offspring:
// a 512-wide MLP (four GELU hidden layers, one scalar output) trained online
// on (model encoding, log-fitness) pairs produced by the simulator.
//
// All hidden layers are 512x512 GEMMs in BF16 with FP32 accumulation on the
// gfx950 matrix cores (v_mfma_f32_16x16x32_bf16).  Every GEMM is in "NT" form,
//   C[m][n] = sum_k A[m][k] * Bt[n][k],
// with both operands K-contiguous, so that a tile's rows load as 16-byte
// vectors into LDS and each lane's MFMA fragment (8 consecutive k) is one
// 16-byte LDS read.  Layouts that the backward pass needs in the other
// orientation are written by the producing epilogue (H and H^T, dZ and
// dZ^T; W and W^T by the optimizer), so no separate transpose kernel runs.
//
// Fragment layout of v_mfma_f32_16x16x32_bf16 (lane l of 64):
//   A: row l%16, k = 8*(l/16) + e (e = 0..7)
//   B: column l%16, k = 8*(l/16) + e
//   D: rows 4*(l/16) + r (r = 0..3), column l%16
#include <hip/hip_runtime.h>

#include <algorithm>

#include "grn_mlp.h"

namespace grn_mlp {

typedef __bf16 bf16;
typedef bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));  // 16 staged bytes (a native vector: stays in registers)

constexpr int BM = 128, BN = 128, BK = 32;  // BK: the K granule every GEMM needs

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

__device__ __forceinline__ float gelu_grad(float z) {
  // d/dz [z Phi(z)] = Phi(z) + z phi(z)
  const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * __expf(-0.5f * z * z);
  return cdf + z * pdf;
}

// C = A . Bt^T for M % TM == 0, N % 128 == 0, K % BKT == 0 (checked by the
// host).  256 threads: 4 waves in a 2 x 2 arrangement, each wave a
// (TM / 2) x 64 block of 16 x 16 accumulators.  TM = 128, 64 or 32 (with
// BKT = 128): the launcher picks the tallest tile that still gives each CU
// two workgroups.  BKT: the K depth
// of one LDS tile: 128 (64 for 128-row tiles, which would need 68 KB of LDS
// at 128) where K allows it -- several MFMA k-steps between two barriers
// and one global-load wait -- else 32.
template <int EPI, int TM, int BKT>
__global__ __launch_bounds__(256) void gemm_nt_kernel(MlpGemm g) {
  constexpr int WI = TM / 32;       // 16-row accumulator blocks per wave
  constexpr int STRIDE = BKT + 8;   // bf16 per LDS tile row: 16 bytes of padding
  constexpr int VR = BKT / 8;                            // 16-byte vectors per tile row
  constexpr int NA = TM * VR / 256, NB = BN * VR / 256;  // vectors per thread of A and of Bt
  static_assert(NA >= 1 && NB >= 1 && VR % NA == 0 && VR % NB == 0, "staging map");
  __shared__ __attribute__((aligned(16))) bf16 As[TM * STRIDE];
  __shared__ __attribute__((aligned(16))) bf16 Bs[BN * STRIDE];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, wm = w >> 1, wn = w & 1;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * BN;
  const int K = g.K;
  // this workgroup's K slice (the whole of K unless split)
  const int ks = (EPI == EPI_F32 && g.ksplit > 1) ? g.ksplit : 1;
  const int kbeg = blockIdx.z * (K / ks), kend = kbeg + K / ks;

  // thread t stages NA (NB) consecutive 16-byte vectors of one tile row of A
  // (Bt): VR / NA threads share a row, so one pointer and immediate offsets
  // serve all of a thread's loads
  const int arow = t / (VR / NA), acol = 8 * NA * (t % (VR / NA));
  const int brow = t / (VR / NB), bcol = 8 * NB * (t % (VR / NB));
  const u32x4 *pa = (const u32x4 *)(g.A + (size_t)(m0 + arow) * K + kbeg + acol);
  const u32x4 *pb = (const u32x4 *)(g.Bt + (size_t)(n0 + brow) * K + kbeg + bcol);
  u32x4 ra[NA], rb[NB];
#pragma unroll
  for (int q = 0; q < NA; ++q) ra[q] = pa[q];
#pragma unroll
  for (int q = 0; q < NB; ++q) rb[q] = pb[q];

  f32x4 acc[WI][4];
#pragma unroll
  for (int i = 0; i < WI; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  const int frow = lane & 15, fk = 8 * (lane >> 4);
  for (int k0 = kbeg; k0 < kend; k0 += BKT) {
    __syncthreads();  // the previous tile's fragment reads are done
#pragma unroll
    for (int q = 0; q < NA; ++q) ((u32x4 *)(As + arow * STRIDE + acol))[q] = ra[q];
#pragma unroll
    for (int q = 0; q < NB; ++q) ((u32x4 *)(Bs + brow * STRIDE + bcol))[q] = rb[q];
    __syncthreads();
    if (k0 + BKT < kend) {  // prefetch the next K slice while this one multiplies
      pa += VR;
      pb += VR;
#pragma unroll
      for (int q = 0; q < NA; ++q) ra[q] = pa[q];
#pragma unroll
      for (int q = 0; q < NB; ++q) rb[q] = pb[q];
    }
#pragma unroll
    for (int kk = 0; kk < BKT; kk += 32) {
      bf16x8 af[WI], bfr[4];
#pragma unroll
      for (int i = 0; i < WI; ++i)
        af[i] = *(const bf16x8 *)(As + ((TM / 2) * wm + 16 * i + frow) * STRIDE + kk + fk);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        bfr[j] = *(const bf16x8 *)(Bs + (64 * wn + 16 * j + frow) * STRIDE + kk + fk);
#pragma unroll
      for (int i = 0; i < WI; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    }
  }

  const int M = g.M, N = g.N;
  float *Cout = ks > 1 ? g.Cpart + (size_t)blockIdx.z * M * N : g.C;
#pragma unroll
  for (int i = 0; i < WI; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int mb = m0 + (TM / 2) * wm + 16 * i + 4 * (lane >> 4);  // first of this lane's 4 rows
      const int n = n0 + 64 * wn + 16 * j + frow;
      float v[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const size_t o = (size_t)(mb + r) * N + n;
        float x = acc[i][j][r];
        if constexpr (EPI == EPI_FWD) {
          x += g.bias[n];
          if (g.C) g.C[o] = x;  // pre-activation, kept for the backward pass
          x = gelu(x);
          g.H[o] = (bf16)x;
        } else if constexpr (EPI == EPI_BWD) {
          x *= gelu_grad(g.Zprev[o]);
          g.H[o] = (bf16)x;
        } else {
          Cout[o] = x;
        }
        v[r] = x;
      }
      if constexpr (EPI != EPI_F32) {
        if (g.Ht) {  // the transposed copy: 4 consecutive rows -> one 8-byte store
          bf16x4 q = {(bf16)v[0], (bf16)v[1], (bf16)v[2], (bf16)v[3]};
          *(bf16x4 *)(g.Ht + (size_t)n * M + mb) = q;
        }
      }
    }
}

// C = the sum of the ks partial products in slice order, four floats per
// thread.
__global__ __launch_bounds__(256) void ksum_kernel(const float4 *__restrict__ part, long n4, int ks,
                                                   float4 *__restrict__ C) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (long)gridDim.x * blockDim.x) {
    float4 s = part[i];
    for (int z = 1; z < ks; ++z) {
      const float4 v = part[(size_t)z * n4 + i];
      s.x += v.x;
      s.y += v.y;
      s.z += v.z;
      s.w += v.w;
    }
    C[i] = s;
  }
}

// Output layer, one wave per row: y[b] = H[b] . w + b0 (512 inputs).
__global__ __launch_bounds__(256) void out_forward_kernel(const bf16 *__restrict__ H, const bf16 *__restrict__ w,
                                                          const float *__restrict__ b0, float *__restrict__ y,
                                                          int rows) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= rows) return;
  const bf16x8 h = *(const bf16x8 *)(H + (size_t)row * MLP_WIDTH + 8 * lane);
  const bf16x8 ww = *(const bf16x8 *)(w + 8 * lane);
  float s = 0.f;
#pragma unroll
  for (int e = 0; e < 8; ++e) s += (float)h[e] * (float)ww[e];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) y[row] = s + b0[0];
}

// Loss gradient at the output and the last hidden layer's dZ:
//   dy[b] = mask[b] * 2 (y[b] - t[b]) / n_valid,
//   dZ[b][k] = dy[b] * w[k] * gelu'(Z[b][k]), written as dZ and dZ^T.
// pair_w > 0: rows 2i and 2i + 1 are a pair (a child and its parent), and
// the loss adds pair_w * ((y_2i - y_2i+1) - (t_2i - t_2i+1))^2 per pair, so
// dy[b] gains 2 pair_w e_b / n_valid with e_b the pair's difference error
// seen from row b (the partner's is -e_b).
// Rows per workgroup of the kernels that also write a transposed copy: the
// workgroup's rows pass through an LDS tile [column][row], so each column of
// the copy is written as GR consecutive values (64 bytes) instead of one
// 2-byte store per value GR columns apart.
constexpr int GR = 32;
typedef bf16 TransposeTile[MLP_WIDTH][GR + 2];  // +2: the rows of two columns start in different banks

// Writes rows r0 .. r0 + GR - 1 (those below `rows`) of the tile into the
// transposed matrix Xt[MLP_WIDTH][rows]: four rows (8 bytes) per store where
// they are aligned, single values otherwise.
__device__ __forceinline__ void store_transposed(const TransposeTile &tile, int r0, int rows, bf16 *__restrict__ Xt) {
  const int nr = rows - r0 < GR ? rows - r0 : GR;
  const bool vec = (rows & 3) == 0;
  for (int q = threadIdx.x; q < MLP_WIDTH * (GR / 4); q += blockDim.x) {
    const int c = q / (GR / 4), r = 4 * (q % (GR / 4));
    bf16 *dst = Xt + (size_t)c * rows + r0 + r;
    if (vec && r + 4 <= nr) {
      const bf16x4 v = {tile[c][r], tile[c][r + 1], tile[c][r + 2], tile[c][r + 3]};
      *(bf16x4 *)dst = v;
    } else {
      for (int e = 0; e < 4 && r + e < nr; ++e) dst[e] = tile[c][r + e];
    }
  }
}

__global__ __launch_bounds__(256) void out_backward_kernel(const float *__restrict__ y, const float *__restrict__ tgt,
                                                           const float *__restrict__ mask, const bf16 *__restrict__ w,
                                                           const float *__restrict__ Z, float inv_n, float pair_w,
                                                           int rows, float *__restrict__ dy, float *__restrict__ sq,
                                                           bf16 *__restrict__ dZ, bf16 *__restrict__ dZt) {
  __shared__ TransposeTile tile;
  const int r0 = blockIdx.x * GR, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const bf16x8 ww = *(const bf16x8 *)(w + 8 * lane);
  for (int i = wave; i < GR && r0 + i < rows; i += 4) {  // one wave per row
    const int row = r0 + i;
    const float e = mask[row] * (y[row] - tgt[row]);
    float ep = 0.f;
    if (pair_w > 0.f) {
      const int q = row ^ 1;
      ep = mask[row] * mask[q] * ((y[row] - y[q]) - (tgt[row] - tgt[q]));
    }
    const float d = 2.f * (e + pair_w * ep) * inv_n;
    if (lane == 0) {
      dy[row] = d;
      sq[row] = e * e + 0.5f * pair_w * ep * ep;  // each pair's term is shared by its two rows
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int col = 8 * lane + k;
      const bf16 v = (bf16)(d * (float)ww[k] * gelu_grad(Z[(size_t)row * MLP_WIDTH + col]));
      dZ[(size_t)row * MLP_WIDTH + col] = v;
      tile[col][i] = v;
    }
  }
  __syncthreads();
  store_transposed(tile, r0, rows, dZt);
}

// out[r] = sum_c X[r][c] * (v ? v[c] : 1) over a bf16 matrix with `cols`
// columns, one workgroup per row, fixed summation order (deterministic).
__global__ __launch_bounds__(256) void row_dot_kernel(const bf16 *__restrict__ X, const float *__restrict__ v,
                                                      int cols, float *__restrict__ out) {
  __shared__ float part[4];
  const int r = blockIdx.x;
  float s = 0.f;
  for (int c = threadIdx.x; c < cols; c += 256) s += (float)X[(size_t)r * cols + c] * (v ? v[c] : 1.f);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[r] = (part[0] + part[1]) + (part[2] + part[3]);
}

// Sum of a float vector into out[0] (one workgroup, fixed order).
__global__ __launch_bounds__(256) void sum_kernel(const float *__restrict__ x, int n, float *__restrict__ out) {
  __shared__ float part[4];
  float s = 0.f;
  for (int i = threadIdx.x; i < n; i += 256) s += x[i];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = (part[0] + part[1]) + (part[2] + part[3]);
}

// Adam over the flat FP32 parameter vector, then the BF16 copies the GEMMs
// read: every hidden weight matrix as W^T (out x in, the forward operand) and
// W (in x out, the backward operand), and the output weights.
__global__ __launch_bounds__(256) void adam_kernel(MlpAdam a) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const float g = a.grad[i];
  const float m = a.beta1 * a.m[i] + (1.f - a.beta1) * g;
  const float v = a.beta2 * a.v[i] + (1.f - a.beta2) * g * g;
  a.m[i] = m;
  a.v[i] = v;
  const float p = a.p[i] - a.lr_t * m / (sqrtf(v) + a.eps_t);
  a.p[i] = p;
  const long mat = (long)MLP_HIDDEN_LAYERS * MLP_WIDTH * MLP_WIDTH;
  if (i < mat) {
    const long l = i / (MLP_WIDTH * MLP_WIDTH), e = i % (MLP_WIDTH * MLP_WIDTH);
    const long o = e / MLP_WIDTH, in = e % MLP_WIDTH;
    a.wt[i] = (bf16)p;
    a.w[l * MLP_WIDTH * MLP_WIDTH + in * MLP_WIDTH + o] = (bf16)p;
  } else if (i >= a.out_w_offset && i < a.out_w_offset + MLP_WIDTH) {
    a.w_out[i - a.out_w_offset] = (bf16)p;
  }
}

// Training batch: rows idx[b] of the sample store (features and targets).
// Targets are standardised: t = (target - mu) * inv_sigma.
__global__ __launch_bounds__(256) void gather_kernel(const bf16 *__restrict__ feat, const float *__restrict__ target,
                                                     const int *__restrict__ idx, int rows, float mu, float inv_sigma,
                                                     bf16 *__restrict__ X, bf16 *__restrict__ Xt,
                                                     float *__restrict__ t) {
  __shared__ TransposeTile tile;
  const int r0 = blockIdx.x * GR, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int i = wave; i < GR && r0 + i < rows; i += 4) {  // one wave per row
    const int row = r0 + i;
    const int src = idx[row];
    const bf16x8 x = *(const bf16x8 *)(feat + (size_t)src * MLP_WIDTH + 8 * lane);
    *(bf16x8 *)(X + (size_t)row * MLP_WIDTH + 8 * lane) = x;
#pragma unroll
    for (int k = 0; k < 8; ++k) tile[8 * lane + k][i] = x[k];
    if (lane == 0) t[row] = (target[src] - mu) * inv_sigma;
  }
  __syncthreads();
  store_transposed(tile, r0, rows, Xt);
}

}  // namespace grn_mlp

using namespace grn_mlp;

template <int TM, int BKT>
static void launch_gemm(int epi, const MlpGemm &g, dim3 grid, hipStream_t s) {
  const dim3 block(256);
  if (epi == EPI_FWD)
    hipLaunchKernelGGL((gemm_nt_kernel<EPI_FWD, TM, BKT>), grid, block, 0, s, g);
  else if (epi == EPI_BWD)
    hipLaunchKernelGGL((gemm_nt_kernel<EPI_BWD, TM, BKT>), grid, block, 0, s, g);
  else
    hipLaunchKernelGGL((gemm_nt_kernel<EPI_F32, TM, BKT>), grid, block, 0, s, g);
}

hipError_t grn_mlp_gemm(int epi, const MlpGemm &g, hipStream_t s) {
  const int ks = (epi == EPI_F32 && g.ksplit > 1) ? g.ksplit : 1;
  if (ks > 1 && (!g.Cpart || g.K % (BK * ks))) return hipErrorInvalidValue;
  // deep K tiles where every workgroup's K slice is a multiple of 128
  const bool deep = (g.K / ks) % 128 == 0;
  // the tallest row tile that still gives at least two workgroups per CU
  // (512; one per CU leaves each global-load wait unhidden), down to 64 rows
  // (32 with deep tiles); a K-split GEMM keeps 128-row tiles (measured: its
  // 16 slices already give each CU a workgroup, and 64-row tiles were 10 %
  // slower)
  auto tiles = [&](int tm) { return (long)(g.N / BN) * (g.M / tm) * ks; };
  int tm = BM;
  if (ks == 1 && tiles(BM) < 512 && g.M % 64 == 0) tm = 64;
  if (tm == 64 && deep && tiles(64) < 512 && g.M % 32 == 0) tm = 32;
  const dim3 grid(g.N / BN, g.M / tm, ks);
  if (tm == 32)
    launch_gemm<32, 128>(epi, g, grid, s);
  else if (tm == 64)
    deep ? launch_gemm<64, 128>(epi, g, grid, s) : launch_gemm<64, BK>(epi, g, grid, s);
  else
    deep ? launch_gemm<BM, 64>(epi, g, grid, s) : launch_gemm<BM, BK>(epi, g, grid, s);
  if (ks > 1) {
    const long n4 = (long)g.M * g.N / 4;
    hipLaunchKernelGGL(ksum_kernel, dim3((unsigned)std::min<long>((n4 + 255) / 256, 1024)), dim3(256), 0, s,
                       (const float4 *)g.Cpart, n4, ks, (float4 *)g.C);
  }
  return hipGetLastError();
}

hipError_t grn_mlp_out_forward(const __bf16 *H, const __bf16 *w, const float *b0, float *y, int rows, hipStream_t s) {
  hipLaunchKernelGGL(out_forward_kernel, dim3((rows + 3) / 4), dim3(256), 0, s, H, w, b0, y, rows);
  return hipGetLastError();
}

hipError_t grn_mlp_out_backward(const float *y, const float *t, const float *mask, const __bf16 *w, const float *Z,
                                float inv_n, float pair_w, int rows, float *dy, float *sq, __bf16 *dZ, __bf16 *dZt,
                                hipStream_t s) {
  if (pair_w > 0.f && rows % 2) return hipErrorInvalidValue;
  hipLaunchKernelGGL(out_backward_kernel, dim3((rows + GR - 1) / GR), dim3(256), 0, s, y, t, mask, w, Z, inv_n, pair_w,
                     rows, dy, sq, dZ, dZt);
  return hipGetLastError();
}

hipError_t grn_mlp_row_dot(const __bf16 *X, const float *v, int rows, int cols, float *out, hipStream_t s) {
  hipLaunchKernelGGL(row_dot_kernel, dim3(rows), dim3(256), 0, s, X, v, cols, out);
  return hipGetLastError();
}

hipError_t grn_mlp_sum(const float *x, int n, float *out, hipStream_t s) {
  hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, s, x, n, out);
  return hipGetLastError();
}

hipError_t grn_mlp_adam(const MlpAdam &a, hipStream_t s) {
  hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t grn_mlp_gather(const __bf16 *feat, const float *target, const int *idx, int rows, float mu,
                          float inv_sigma, __bf16 *X, __bf16 *Xt, float *t, hipStream_t s) {
  hipLaunchKernelGGL(gather_kernel, dim3((rows + GR - 1) / GR), dim3(256), 0, s, feat, target, idx, rows, mu, inv_sigma, X,
                     Xt, t);
  return hipGetLastError();
}
