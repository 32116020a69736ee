// grn_exact.hip -- the FP64 field kernel: the reference's double-precision
// simulation (Src/Simulator/modelsimulator.cpp:205-280, simulator.cpp:54-63,
// Src/Search/evaluatormorphogens.cpp:48-76) for runs whose result hangs on
// the convergence measure.
//
// With an early-stop threshold (minConcChange >= 0) the number of Euler steps
// is an integer decision on the last step's max |dc/dt|, and in REF fitness
// that same value is added to the error when the run does not converge.
// Near convergence the rate is a small difference of large terms, so the
// FP32 kernels' rounding of the Hill terms (~1e-7 of lim*h) is a relative
// error of ~1e-4 in it: enough to move the stop by a step or the penalty in
// its fourth digit.  This kernel keeps the state and every operation in
// double precision, as the reference does, so the stop step and the penalty
// are the reference's.
//
// Layout: one workgroup per model (persistent, the same work queues as the
// other kernels), 256 threads dealt over the cells; the state lives in the
// workgroup's device-memory slot as two plane sets A (step start) and B
// (step end), [G][H][W] doubles each, so any grid size runs.  Each thread
// keeps its cell's genes in registers and their log2 in its own column of an
// LDS array (the links index it with the regulator).  One barrier per step:
// after it every thread has finished reading A and writing B, the sets swap,
// and the per-wave max changes (double-buffered by step parity) are read.
// Hill terms use the exponent form of grn_program.h in double:
// h = 1/(1 + 2^T), T = max(min_OR t, max_AND t), t = a*log2(x) + b.
#include <hip/hip_runtime.h>

#include "grn_kernels.h"
#include "grn_program.h"

namespace grn_exact {

constexpr double kMaxD = 1.7976931348623157e308;

template <int NT>
struct Reduce {
  double wmax[2][NT / 64];  // per-wave max |R|, by step parity
  double wsum[NT / 64];     // per-wave fitness partial sums
};

// MathAlgo::max (mathalgo.h:213-216): b if a < b, else a.
__device__ __forceinline__ double ref_max(double a, double b) { return a < b ? b : a; }

template <int G, int kNT>
__device__ __forceinline__ void run_model(const GrnLaunch &a, const GrnProgram *__restrict__ progs,
                                          const GrnDataset *__restrict__ datasets, int m, int pi,
                                          double *lgs, Reduce<kNT> &red) {
#pragma clang fp contract(off)
  constexpr int kWaves = kNT / 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const GrnProgram *P = progs + pi;
  const GrnProgram64 *Q = a.progs64 + pi;
  const GrnDataset *D = datasets + P->dataset;
  const int W = D->W, H = D->H, N = W * H;
  double *A = reinterpret_cast<double *>(a.scratch + (size_t)blockIdx.x * a.slot_floats);
  double *B = A + (size_t)G * N;
  const double dt = Q->dt, min_conc = D->min_conc_d, mcc = D->min_conc_change_d;
  const int n_exp = D->n_exp, n_in = D->n_in, n_tgt = D->n_tgt;
  const int S = (a.step_cap > 0 && a.step_cap < D->steps) ? a.step_cap : D->steps;
  const double max_err = a.max_err[pi];
  const int reg_mask = P->reg_mask, dif_mask = P->dif_mask;
  double *lg = lgs + tid;  // this thread's column: gene g at lg[g * kNT]

  double err = 0.0, chg_acc = 0.0;
  int e = 0, steps_total = 0;
  while (e < n_exp && err <= max_err) {
    // Initial state: input planes for genes < nIn, zero elsewhere
    // (simulator.cpp:35-52).
    const double *init = D->init + (size_t)e * n_in * N;
    for (int i = tid; i < G * N; i += kNT) A[i] = i < n_in * N ? init[i] : 0.0;
    __syncthreads();

    double chg = 1.0;  // simulator.cpp:55
    int s = 0;
    while (s < S && chg > mcc) {
      double cmax = 0.0;
      for (int cell = tid; cell < N; cell += kNT) {
        const int x = cell % W, y = cell / W;
        double c[G];
#pragma unroll
        for (int g = 0; g < G; ++g) c[g] = A[(size_t)g * N + cell];
#pragma unroll
        for (int g = 0; g < G; ++g)
          if ((reg_mask >> g) & 1) lg[g * kNT] = ref_max(log2(c[g]), -kMaxD);
#pragma unroll
        for (int p = 0; p < G; ++p) {
          const GrnGene &gp = P->gene[p];
          const GrnGene64 &gq = Q->gene[p];
          // the fold min(max_OR H, min_AND H) of modelsimulator.cpp:156-192
          double h = 0.0;
          const int lb = gp.link_begin, nor = gp.n_or, nand = gp.n_and;
          if (nor + nand > 0) {
            double T = 0.0;
            if (nor > 0) {
              T = Q->link[lb].a * lg[P->link[lb].r * kNT] + Q->link[lb].b;
              for (int l = lb + 1; l < lb + nor; ++l)
                T = fmin(T, Q->link[l].a * lg[P->link[l].r * kNT] + Q->link[l].b);
            }
            if (nand > 0) {
              const int l0 = lb + nor;
              double U = Q->link[l0].a * lg[P->link[l0].r * kNT] + Q->link[l0].b;
              for (int l = l0 + 1; l < l0 + nand; ++l)
                U = fmax(U, Q->link[l].a * lg[P->link[l].r * kNT] + Q->link[l].b);
              T = nor > 0 ? fmax(T, U) : U;
            }
            h = 1.0 / (1.0 + exp2(T));
          }
          // modelsimulator.cpp:245-275: rate, guarded diffusion from the
          // step-start field, update only where the rate is nonzero, minConc
          // clamp, max |rate|
          double R = gq.lim * h - gq.deg * c[p];
          if ((dif_mask >> p) & 1) {
            const double *pl = A + (size_t)p * N;
            double sum = 0.0, n = 0.0;
            if (x > 0) { sum += pl[cell - 1]; n += 1.0; }
            if (x < W - 1) { sum += pl[cell + 1]; n += 1.0; }
            if (y > 0) { sum += pl[cell - W]; n += 1.0; }
            if (y < H - 1) { sum += pl[cell + W]; n += 1.0; }
            if (sum > 0 || (n > 0 && c[p] > 0)) R += gq.dif * (sum - n * c[p]);
          }
          double v = c[p];
          if (R != 0.0) {
            v = c[p] + dt * R;
            if (v < min_conc) v = 0.0;
            cmax = ref_max(cmax, fabs(R));
          }
          B[(size_t)p * N + cell] = v;
        }
      }
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) cmax = ref_max(cmax, __shfl_xor(cmax, off));
      if (lane == 0) red.wmax[s & 1][wave] = cmax;
      __syncthreads();  // every read of A and write of B of this step is done
      chg = 0.0;
#pragma unroll
      for (int w = 0; w < kWaves; ++w) chg = ref_max(chg, red.wmax[s & 1][w]);
      double *t = A;
      A = B;
      B = t;
      ++s;
    }
    steps_total += s;
    chg_acc += chg;

    // REF: sum of log(1 + (|c - T| - thr)) over positive terms, / (W*H)
    // (simstate.cpp:135-158); MSE: mean of (c - T)^2 over cells and targets.
    const bool mse = D->kind == 1;
    const double *tgt = D->target + (size_t)e * n_tgt * N;
    const double *fin = A + (size_t)n_in * N;
    double acc = 0.0;
    for (int i = tid; i < n_tgt * N; i += kNT) {
      const double d = fin[i] - tgt[i];
      if (mse) {
        acc += d * d;
      } else {
        const double u = fabs(d) - D->thr;
        if (u > 0) acc += log(1.0 + u);
      }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) red.wsum[wave] = acc;
    __syncthreads();
    double tot = 0.0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) tot += red.wsum[w];
    double e_err;
    if (mse) {
      e_err = tot / ((double)N * n_tgt);
    } else {
      e_err = tot / N;
      if (chg_acc > mcc) e_err += chg_acc - mcc;  // evaluatormorphogens.cpp:62-63
    }
    err += e_err / n_exp;
    if (a.fields && e == a.field_exp) {
      double *fo = a.fields + (size_t)m * a.field_stride;
      for (int i = tid; i < G * N; i += kNT) fo[i] = A[i];
    }
    ++e;
    __syncthreads();  // wsum and the final state are read before the next experiment
  }

  if (tid == 0) {
    GrnModelOut o;
    o.err = err;
    o.change = chg_acc;
    o.steps = steps_total;
    o.exps = e;
    o.done = 1;
    o.pad = 0;
    a.out[pi] = o;
  }
}

template <int G, int GHI, int NT>
__device__ __forceinline__ void dispatch(int g, const GrnLaunch &a, const GrnProgram *__restrict__ progs,
                                         const GrnDataset *__restrict__ datasets, int m, int pi, double *lgs,
                                         Reduce<NT> &red) {
  if constexpr (G <= GHI) {
    if (g == G)
      run_model<G, NT>(a, progs, datasets, m, pi, lgs, red);
    else
      dispatch<G + 1, GHI, NT>(g, a, progs, datasets, m, pi, lgs, red);
  }
}

template <int GLO, int GHI, int kNT>
__global__ __launch_bounds__(kNT) void exact_kernel(GrnLaunch a, const GrnProgram *__restrict__ progs,
                                                    const GrnDataset *__restrict__ datasets,
                                                    const int *__restrict__ order) {
  __shared__ double s_lg[GHI * kNT];
  __shared__ Reduce<kNT> s_red;
  __shared__ int s_model;
  GrnQueue *q = a.queue;
  const int n = __builtin_amdgcn_readfirstlane(q->n);
  const int *ord = order + __builtin_amdgcn_readfirstlane(q->begin);
  for (;;) {
    if (threadIdx.x == 0) s_model = atomicAdd(&q->head, 1);
    __syncthreads();
    const int m = __builtin_amdgcn_readfirstlane(s_model);
    __syncthreads();
    if (m >= n) break;
    const int pi = __builtin_amdgcn_readfirstlane(ord[m]);
    const int g = __builtin_amdgcn_readfirstlane(progs[pi].G);
    dispatch<GLO, GHI, kNT>(g, a, progs, datasets, m, pi, s_lg, s_red);
    __syncthreads();
  }
}

}  // namespace grn_exact

// wc = -2 marks the FP64 kernel; it applies the diffusion guard always.
// Two gene ranges: the 1-8-gene bodies need fewer registers (152 against
// 204 VGPRs), so three workgroups share a CU instead of two.
static const GrnKernelEntry kTable[] = {
    {-2, 1, 256, 1, 8, true, (const void *)&grn_exact::exact_kernel<1, 8, 256>, 0},
    {-2, 1, 256, 9, 16, true, (const void *)&grn_exact::exact_kernel<9, 16, 256>, 0},
};

extern "C" const GrnKernelEntry *grn_field_table_exact(int *n) {
  *n = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}
