// grn_field.hpp -- the on-chip field kernel (included by the per-width
// translation units grn_field_w*.hip, which pick the instantiations).
//
// One workgroup simulates one model at a time; a persistent grid pulls models
// from an atomic work queue sorted costliest-first.  Each model's whole field
// stays on chip for all Euler steps:
//   * every thread owns a vertical strip of CPT cells (column x, rows
//     y0..y0+CPT-1) and keeps all G genes of those cells in registers as a
//     compensated pair (hi, lo) of FP32 values;
//   * every gene's plane is mirrored in LDS with a zero border, from which a
//     step reads the four neighbours (two ds_read2 per gene and cell), so
//     only the owning thread's registers hold the state;
//   * Hill regulation is evaluated in exponent form (grn_program.h): one log2
//     per gene, one FMA + min/max per link, one exp2 and one reciprocal per
//     gene; each pass over a gene's links serves a block of CB cells of the
//     strip at once, which amortises the per-link scalar work (load, register
//     index set-up, loop control) -- the scalar pipe, not the vector ALU,
//     limits a one-cell-per-pass loop;
//   * the per-step max |dc/dt| (the reference's convergence test) and the
//     final fitness are reduced in the same kernel.
// The gene count G is a template parameter: the kernel switches once per
// model to the body compiled for that G, so gene loops are fully unrolled
// with no per-gene conditions.
//
// Semantics follow the reference CPU simulator: Src/Simulator/
// modelsimulator.cpp:205-280 (step), simulator.cpp:54-63 (stop rule),
// simstate.cpp:135-158 (log error), Src/Search/evaluatormorphogens.cpp:48-76
// (experiment loop, penalty, maxError cut).  Like the CPU path -- and unlike
// the reference CUDA kernel -- every cell of the W x H grid is simulated and
// scored.
#pragma once
#include <hip/hip_runtime.h>

#include "grn_kernels.h"
#include "grn_program.h"

namespace grn_field {

// Per-cell log2 values, indexed by a uniform regulator index.  Built with
// -amdgpu-use-divergent-register-indexing so that every width is read with
// one indexed register move (s_set_gpr_idx) rather than a select chain.
template <int G>
struct LgVec {
  typedef float T __attribute__((ext_vector_type(G <= 4 ? 4 : (G <= 8 ? 8 : 16))));
};

__device__ __forceinline__ float hw_log2(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ float hw_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float hw_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

constexpr float kMaxF = 3.402823466e38f;

// Opaque values: the compiler cannot prove them loop-invariant, so loads and
// address arithmetic derived from them stay where they are written instead of
// being hoisted into long-lived registers.  opaque_zero() is uniform (it goes
// through readfirstlane), so a pointer offset by it keeps its scalar loads.
__device__ __forceinline__ int opaque_zero() {
  int z = 0;
  asm volatile("" : "+v"(z));
  return __builtin_amdgcn_readfirstlane(z);
}
__device__ __forceinline__ int opaque_v(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

struct Shared {
  float *lds;      // planes (dynamic LDS)
  float *wmax;     // per-wave max change
  double *wsum;    // per-wave fitness partial sums
};

// Everything one model (or one tile of a model) needs; G, CPT, WC
// compile-time.  TILED: the registers hold an extended tile of a larger grid
// (grn_tiled.hpp): cells outside the grid get a zero time step so they stay
// zero, and only the tile's interior cells count towards the max change.
template <int WC, int G, int CPT, int NTMAX, bool GUARD, bool TILED = false>
struct Model {
  const GrnProgram *P;
  int W, H, plane;
  float dt, min_conc;
  bool clamp_any_r;      // the R != 0 test of the clamp is moot (min_conc <= 0, field >= 0)
  bool active;
  int x, y0, ubase;      // ubase: LDS index of the cell north-west of cell k=0
  float nf[TILED ? 1 : CPT];  // neighbour count per cell (on-chip kernels)
  // TILED: the per-cell constants of the current window packed into two
  // registers instead of 3 CPT: nfb holds each cell's neighbour count in 3
  // bits (cell k at bit 3k); cellb bit k = the cell lies in the grid (time
  // step dt, else 0), bit 16 + k = its |R| counts towards the max change.
  unsigned nfb, cellb;
  // TILED: the unpacked constants of the cells of the block being stepped
  // (written at the start of each cell block, live only within it)
  float blk_nf[TILED ? CPT : 1], blk_dt[TILED ? CPT : 1];
  float hi[G][CPT];
  float lo[G][CPT];
  Shared sh;
  static_assert(!TILED || CPT <= 10, "the packed per-cell constants hold at most 10 cells");

  __device__ __forceinline__ int pc() const { return WC ? WC + 2 : W + 2; }
  __device__ __forceinline__ float nfv(int k) const {
    if constexpr (TILED)
      return blk_nf[k];
    else
      return nf[k];
  }
  __device__ __forceinline__ float dtv(int k) const {
    if constexpr (TILED)
      return blk_dt[k];
    else
      return dt;
  }
  // TILED: unpacks the constants of cells k0 .. k0 + n - 1 once per cell
  // block, so the genes' updates read them instead of re-deriving them.
  template <int n>
  __device__ __forceinline__ void unpack_block(int k0) {
    if constexpr (TILED) {
      asm volatile("" : "+v"(nfb), "+v"(cellb));
#pragma unroll
      for (int j = 0; j < n; ++j) {
        const int k = k0 + j;
        blk_nf[k] = (float)((nfb >> (3 * k)) & 7u);
        blk_dt[k] = ((cellb >> k) & 1u) ? dt : 0.f;
        asm volatile("" : "+v"(blk_nf[k]), "+v"(blk_dt[k]));
      }
    }
  }
  __device__ __forceinline__ bool trkv(int k) const {
    if constexpr (TILED)
      return ((cellb >> (16 + k)) & 1u) != 0;
    else
      return true;
  }

  // Only diffusing genes are read back from the LDS mirror.
  __device__ __forceinline__ void store_planes() {
    if (!active) return;
    const int dm = P->dif_mask;
    float *pl = sh.lds + ubase + pc() + 1;
#pragma unroll
    for (int p = 0; p < G; ++p)
      if (dm & (1 << p))
#pragma unroll
        for (int k = 0; k < CPT; ++k) pl[p * plane + k * pc()] = hi[p][k];
  }

  // One Euler step of every owned cell, processed cell by cell down the
  // strip.  With TRACK the workgroup-wide max |R| of the step is returned.
  // Cells per link pass: every link iteration serves CB cells, which
  // amortises its scalar work (load, index set-up, loop control); CB is the
  // largest divisor of CPT whose log2 vectors still fit the register budget
  // of NTMAX threads next to the (hi, lo) state.
  static constexpr int lg_width() { return G <= 4 ? 4 : (G <= 8 ? 8 : 16); }
  static constexpr int budget() { return 131072 / NTMAX < 512 ? 131072 / NTMAX : 512; }
  static constexpr bool fits(int cb) {
    return 2 * G * CPT + (TILED ? 3 * CPT + 32 : 24) + lg_width() * cb + 4 * cb <= budget();
  }
  static constexpr int cell_block() {
    int best = 1;
    for (int cb = 1; cb <= CPT; ++cb)
      if (CPT % cb == 0 && fits(cb)) best = cb;
    return best;
  }

  // R = lim*h - deg*c (+ dif * the 5-point Laplacian from the LDS mirror
  // when DIF) and the compensated Euler update of gene p for the cells
  // k0 .. k0 + CB - 1.
  template <bool TRACK, bool FAST, bool DIF, int CB>
  __device__ __forceinline__ void update_cells_n(int p, const GrnGene &gp, const float *hv, int k0, float *cmax) {
    const float *pl = sh.lds + p * plane + ubase;
#pragma unroll
    for (int j = 0; j < CB; ++j) {
      const int k = k0 + j;
      const float c = hi[p][k];
      const float h = hv[j];
      float R = __builtin_fmaf(gp.lim, h, -gp.deg * c);
      if (DIF) {
        const float left = pl[(k + 1) * pc()];
        const float right = pl[(k + 1) * pc() + 2];
        const float up = pl[k * pc() + 1];
        const float down = pl[(k + 2) * pc() + 1];
        const float S = ((left + right) + up) + down;
        float lap = __builtin_fmaf(-nfv(k), c, S);
        if (GUARD) {  // modelsimulator.cpp:255; moot for a nonnegative field
          const bool use = (S > 0.f) || (nfv(k) > 0.f && c > 0.f);
          lap = use ? lap : 0.f;
        }
        R = __builtin_fmaf(gp.dif, lap, R);
      }
      // Compensated Euler update c += dt*R (Fast2Sum on the pair) with
      // the minConc clamp of modelsimulator.cpp:263-275.
      const float yv = __builtin_fmaf(dtv(k), R, lo[p][k]);
      const float sv = c + yv;
      const float nl = yv - (sv - c);
      if (FAST) {
        hi[p][k] = sv;
        lo[p][k] = nl;
      } else {
        const bool clamp = (sv < min_conc) & (clamp_any_r | (R != 0.f));  // branch-free
        hi[p][k] = clamp ? 0.f : sv;
        lo[p][k] = clamp ? 0.f : nl;
      }
      if (TRACK) *cmax = __builtin_fmaxf(*cmax, trkv(k) ? __builtin_fabsf(R) : 0.f);
    }
  }
  template <bool TRACK, bool FAST, bool DIF>
  __device__ __forceinline__ void update_cells(int p, const GrnGene &gp, const float *hv, int k0, float *cmax) {
    update_cells_n<TRACK, FAST, DIF, cell_block()>(p, gp, hv, k0, cmax);
  }
  // The same with the diffusion term only where dif_on (uniform): the four
  // LDS reads and the Laplacian are skipped for a gene that does not diffuse.
  template <bool TRACK, bool FAST>
  __device__ __forceinline__ void update_cells_masked(int p, const GrnGene &gp, const float *hv, int k0, int dif_on,
                                                      float *cmax) {
    constexpr int CB = cell_block();
    float Rv[CB];
#pragma unroll
    for (int j = 0; j < CB; ++j) Rv[j] = __builtin_fmaf(gp.lim, hv[j], -gp.deg * hi[p][k0 + j]);
    if (dif_on) {
      const float *pl = sh.lds + p * plane + ubase;
#pragma unroll
      for (int j = 0; j < CB; ++j) {
        const int k = k0 + j;
        const float c = hi[p][k];
        const float left = pl[(k + 1) * pc()];
        const float right = pl[(k + 1) * pc() + 2];
        const float up = pl[k * pc() + 1];
        const float down = pl[(k + 2) * pc() + 1];
        const float S = ((left + right) + up) + down;
        float lap = __builtin_fmaf(-nfv(k), c, S);
        if (GUARD) {
          const bool use = (S > 0.f) || (nfv(k) > 0.f && c > 0.f);
          lap = use ? lap : 0.f;
        }
        Rv[j] = __builtin_fmaf(gp.dif, lap, Rv[j]);
      }
    }
#pragma unroll
    for (int j = 0; j < CB; ++j) {
      const int k = k0 + j;
      const float c = hi[p][k];
      const float R = Rv[j];
      const float yv = __builtin_fmaf(dtv(k), R, lo[p][k]);
      const float sv = c + yv;
      const float nl = yv - (sv - c);
      if (FAST) {
        hi[p][k] = sv;
        lo[p][k] = nl;
      } else {
        const bool clamp = (sv < min_conc) & (clamp_any_r | (R != 0.f));
        hi[p][k] = clamp ? 0.f : sv;
        lo[p][k] = clamp ? 0.f : nl;
      }
      if (TRACK) *cmax = __builtin_fmaxf(*cmax, trkv(k) ? __builtin_fabsf(R) : 0.f);
    }
  }

  // FAST: the program guarantees a nonnegative state (no minConc clamp) and
  // no zero Hill exponent (no log2 floor), see GrnProgram::fast.
  // MASKED: some genes do not diffuse (GrnProgram::dif_mask); their
  // diffusion term is skipped by a uniform branch per gene.  Models whose
  // genes all diffuse run the branch-free loop.
  template <bool TRACK, bool FAST, bool MASKED = false>
  __device__ __forceinline__ float step() {
    typedef typename LgVec<G>::T lg_t;
    constexpr int CB = cell_block();
    float cmax = 0.f;

#pragma unroll
    for (int k0 = 0; k0 < CPT; k0 += CB) {
      // Reload the program through an opaque offset in every block, so its
      // constants are not kept in scalar registers across blocks.
      const GrnProgram *Pq = P + opaque_zero();
      // (tiled: the packed per-cell constants are unpacked per block, not
      // hoisted out of the step loop into registers for every cell)
      unpack_block<CB>(k0);
      // log2 of every gene of this block's cells (their step-start values:
      // a cell's update touches only that cell)
      lg_t lg[CB];
#pragma unroll
      for (int j = 0; j < CB; ++j)
#pragma unroll
        for (int g = 0; g < G; ++g) lg[j][g] = FAST ? hw_log2(hi[g][k0 + j]) : __builtin_fmaxf(hw_log2(hi[g][k0 + j]), -kMaxF);
      // On-chip kernels: define the spare lanes and pin each vector as one
      // register tuple, so the indexed reads below use it in place instead of
      // the compiler re-assembling a copy before every gene's link loop.  (The
      // tiled kernels keep more state live and would spill instead.)
      if constexpr (!TILED) {
#pragma unroll
        for (int j = 0; j < CB; ++j) {
#pragma unroll
          for (int g = G; g < lg_width(); ++g) lg[j][g] = 0.f;
          asm volatile("" : "+v"(lg[j]));
        }
      }

#pragma unroll
      for (int p = 0; p < G; ++p) {
        const GrnGene gp = Pq->gene[p];
        // h for the block's cells: the reference's fold min(max_OR H, min_AND H)
        // in exponent form, T = max(min_OR t, max_AND t), h = 1/(1 + 2^T).
        // Each group starts from its first link (no +-inf seeds); a gene
        // without an activator has no stored links and h = 0, without the
        // exponential and reciprocal.  One 16-byte scalar load per link:
        // (r, a, b, pad).
        const float4 *lk = reinterpret_cast<const float4 *>(Pq->link);
        const int lb = gp.link_begin, nor = gp.n_or, nand = gp.n_and;
        float hv[CB];
        if (nor + nand > 0) {
          float T[CB];
          if (nor > 0) {
            const float4 L0 = lk[lb];
            const int r0 = __float_as_int(L0.x);
#pragma unroll
            for (int j = 0; j < CB; ++j) T[j] = __builtin_fmaf(L0.y, lg[j][r0], L0.z);
            for (int l = lb + 1; l < lb + nor; ++l) {
              const float4 L = lk[l];
              const int r = __float_as_int(L.x);
#pragma unroll
              for (int j = 0; j < CB; ++j) T[j] = __builtin_fminf(T[j], __builtin_fmaf(L.y, lg[j][r], L.z));
            }
          }
          if (nand > 0) {
            float U[CB];
            const float4 L0 = lk[lb + nor];
            const int r0 = __float_as_int(L0.x);
#pragma unroll
            for (int j = 0; j < CB; ++j) U[j] = __builtin_fmaf(L0.y, lg[j][r0], L0.z);
            for (int l = lb + nor + 1; l < lb + nor + nand; ++l) {
              const float4 L = lk[l];
              const int r = __float_as_int(L.x);
#pragma unroll
              for (int j = 0; j < CB; ++j) U[j] = __builtin_fmaxf(U[j], __builtin_fmaf(L.y, lg[j][r], L.z));
            }
#pragma unroll
            for (int j = 0; j < CB; ++j) T[j] = nor > 0 ? __builtin_fmaxf(T[j], U[j]) : U[j];
          }
#pragma unroll
          for (int j = 0; j < CB; ++j) hv[j] = hw_rcp(1.0f + hw_exp2(T[j]));
        } else {
#pragma unroll
          for (int j = 0; j < CB; ++j) hv[j] = 0.f;
        }
        if constexpr (MASKED)
          update_cells_masked<TRACK, FAST>(p, gp, hv, k0, (Pq->dif_mask >> p) & 1, &cmax);
        else
          update_cells<TRACK, FAST, true>(p, gp, hv, k0, &cmax);
        // Take this gene's results here so the compiler cannot sink its
        // update below the next genes' link loops (which would keep every
        // gene's constants and partial sums alive at once).
#pragma unroll
        for (int j = 0; j < CB; ++j) asm volatile("" : "+v"(hi[p][k0 + j]), "+v"(lo[p][k0 + j]));
      }
    }

    if (TRACK) {
      if (!active) cmax = 0.f;
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) cmax = __builtin_fmaxf(cmax, __shfl_xor(cmax, off));
    }
    __syncthreads();  // every read of the old planes is done
    store_planes();
    if (TRACK && (threadIdx.x & 63) == 0) sh.wmax[threadIdx.x >> 6] = cmax;
    __syncthreads();  // new planes visible
    float chg = 0.f;
    if (TRACK) {
      const int nw = blockDim.x >> 6;
      for (int w = 0; w < nw; ++w) chg = __builtin_fmaxf(chg, sh.wmax[w]);
    }
    return chg;
  }

  // Simulator::simulate (simulator.cpp:54-63): run while step < S and the
  // last step's max change exceeds minConcChange.  A negative threshold can
  // never stop the loop, so then only the final step needs the reduction.
  template <bool FAST, bool MASKED>
  __device__ __forceinline__ float run_steps(int S, double mcc, int *steps_done) {
    float chg = 1.0f;
    int s = 0;
    if (mcc >= 0.0) {
      while (s < S && (double)chg > mcc) {
        chg = step<true, FAST, MASKED>();
        ++s;
      }
    } else {
      for (; s < S - 1; ++s) step<false, FAST, MASKED>();
      if (S > 0) {
        chg = step<true, FAST, MASKED>();
        ++s;
      }
    }
    *steps_done = s;
    return chg;
  }
};

// Sum of the per-cell fitness terms over this thread's share of the staged
// final planes: REF terms log(1 + (|c - T| - thr)) where positive
// (simstate.cpp:135-158), MSE terms (c - T)^2.  The terms are dealt to the
// first nb threads (the model's own thread count, rounded to whole waves),
// not to the whole workgroup, so the order of the FP64 sum -- and with it
// every bit of the fitness -- does not depend on the workgroup size a batch
// gives the model; threads beyond nb contribute exact zeros.  Kept out of
// line so its double-precision log does not share a register allocation
// with the simulation loop.
__device__ __noinline__ double score_terms(const double *stage, const double *target, int count,
                                           double thr, int mse, int nb) {
  double acc = 0.0;
  if ((int)threadIdx.x >= nb) return acc;
  for (int i = threadIdx.x; i < count; i += nb) {
    const double d = stage[i] - target[i];
    if (mse) {
      acc += d * d;
    } else {
      const double u = __builtin_fabs(d) - thr;
      if (u > 0) acc += log(1.0 + u);
    }
  }
  return acc;
}

// Simulates and scores one model (all experiments) -- the body of the
// reference's simulateModelKernel (modelsimulatordevice.cu:216-251) with the
// CPU semantics of EvaluatorMorphogens::evaluate.
template <int WC, int G, int CPT, int NTMAX, bool GUARD>
__device__ __forceinline__ void run_model(const GrnLaunch &a, const GrnProgram *__restrict__ progs,
                                          const GrnDataset *__restrict__ datasets, int m, int pi,
                                          Shared sh) {
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
  Model<WC, G, CPT, NTMAX, GUARD> f;
  const GrnProgram *P = progs + pi;
  const GrnDataset *D = datasets + P->dataset;
  f.P = P;
  f.sh = sh;
  f.W = WC ? WC : D->W;
  f.H = D->H;
  f.plane = f.pc() * (f.H + 2);
  f.dt = P->dt;
  f.min_conc = D->min_conc;
  f.clamp_any_r = !(D->min_conc > 0.f || !D->nonneg);
  const int W = f.W, H = f.H, N = W * H;
  const int n_active = W * (H / CPT);  // the launcher guarantees H % CPT == 0
  f.active = tid < n_active;
  f.x = f.active ? tid % W : 0;
  f.y0 = f.active ? (tid / W) * CPT : 0;
  f.ubase = f.y0 * f.pc() + f.x;
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    const int y = f.y0 + k;
    f.nf[k] = (float)((f.x > 0) + (f.x < W - 1) + (y > 0) + (y < H - 1));
  }
  const int n_exp = D->n_exp, n_in = D->n_in, n_tgt = D->n_tgt;
  const int S = (a.step_cap > 0 && a.step_cap < D->steps) ? a.step_cap : D->steps;
  const double mcc = D->min_conc_change_d;
  const double max_err = a.max_err[pi];
  const int plane_floats = G * f.plane;
  const int nb = (n_active + 63) & ~63;  // the model's own thread count in whole waves

  double err = 0.0, chg_acc = 0.0;
  int e = 0, steps_total = 0;
  while (e < n_exp && err <= max_err) {
    // Initial state: input planes for genes < nIn, zero elsewhere
    // (simulator.cpp:35-52); each double input is split into hi + lo.
    {
      const int cell = opaque_v(f.y0 * W + f.x);
      const double *ip = D->init + (size_t)e * n_in * N + cell;
#pragma unroll
      for (int g = 0; g < G; ++g)
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
          double v = 0.0;
          if (g < n_in && f.active) v = ip[(size_t)g * N + k * W];
          f.hi[g][k] = (float)v;
          f.lo[g][k] = (float)(v - (double)f.hi[g][k]);
        }
    }
    for (int i = tid; i < plane_floats; i += blockDim.x) sh.lds[i] = 0.f;
    __syncthreads();
    f.store_planes();
    __syncthreads();

    int sdone = 0;
    // a model with non-diffusing genes skips their diffusion terms; a model
    // whose genes all diffuse runs the branch-free loops (the any-width
    // fallback kernels keep the branch-free loops only, to bound code size)
    const bool masked = WC != 0 && P->dif_mask != (1 << G) - 1;
    float chg;
    if (P->fast)
      chg = masked ? f.template run_steps<true, true>(S, mcc, &sdone) : f.template run_steps<true, false>(S, mcc, &sdone);
    else
      chg = masked ? f.template run_steps<false, true>(S, mcc, &sdone) : f.template run_steps<false, false>(S, mcc, &sdone);
    steps_total += sdone;
    chg_acc += (double)chg;

    // Stage the final target planes in LDS as doubles (c = hi + lo), then
    // score them in a plain loop.
    double *stage = (double *)sh.lds;
    if (f.active) {
      const int cell = opaque_v(f.y0 * W + f.x);
#pragma unroll
      for (int g = 0; g < G; ++g)
        if (g >= n_in && g < n_in + n_tgt)
#pragma unroll
          for (int k = 0; k < CPT; ++k)
            stage[(g - n_in) * N + cell + k * W] = (double)f.hi[g][k] + (double)f.lo[g][k];
      if (a.fields && e == a.field_exp) {
        double *fo = a.fields + (size_t)m * a.field_stride + cell;
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
          for (int k = 0; k < CPT; ++k) fo[(size_t)g * N + k * W] = (double)f.hi[g][k] + (double)f.lo[g][k];
      }
    }
    __syncthreads();

    // REF: sum of log(1 + (|c - T| - thr)) over positive terms, / (W*H)
    // (simstate.cpp:135-158).  MSE: mean of (c - T)^2 over cells and targets.
    const bool mse = D->kind == 1;
    double acc = score_terms(stage, D->target + (size_t)e * n_tgt * N, n_tgt * N, D->thr, mse ? 1 : 0, nb);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) sh.wsum[wave] = acc;
    __syncthreads();
    double tot = 0.0;
    for (int w = 0; w < (nb >> 6); ++w) tot += sh.wsum[w];

    double e_err;
    if (mse) {
      e_err = tot / ((double)N * n_tgt);
    } else {
      e_err = tot / N;
      if (chg_acc > mcc) e_err += chg_acc - mcc;  // evaluatormorphogens.cpp:62-63
    }
    err += e_err / n_exp;
    ++e;
    __syncthreads();  // stage and wsum reads done before reuse
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

template <int WC, int CPT, int NTMAX, bool GUARD, int G, int GHI>
__device__ __forceinline__ void dispatch_g(int g, const GrnLaunch &a, const GrnProgram *__restrict__ progs,
                                           const GrnDataset *__restrict__ datasets, int m, int pi,
                                           Shared sh) {
  if constexpr (G <= GHI) {
    if (g == G)
      run_model<WC, G, CPT, NTMAX, GUARD>(a, progs, datasets, m, pi, sh);
    else
      dispatch_g<WC, CPT, NTMAX, GUARD, G + 1, GHI>(g, a, progs, datasets, m, pi, sh);
  }
}

// Persistent kernel: models with GLO <= G <= GHI genes of one grid width.
template <int WC, int CPT, int NTMAX, int GLO, int GHI, bool GUARD>
__global__ __launch_bounds__(NTMAX) void field_kernel(GrnLaunch a, const GrnProgram *__restrict__ progs,
                                                      const GrnDataset *__restrict__ datasets,
                                                      const int *__restrict__ order) {
  extern __shared__ float lds[];
  __shared__ int s_model;
  __shared__ float s_wmax[NTMAX / 64];
  __shared__ double s_wsum[NTMAX / 64];
  Shared sh{lds, s_wmax, s_wsum};
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
    dispatch_g<WC, CPT, NTMAX, GUARD, GLO, GHI>(g, a, progs, datasets, m, pi, sh);
    __syncthreads();
  }
}

}  // namespace grn_field

#define GRN_FIELD_ENTRY(WC, CPT, NTMAX, GLO, GHI, GUARD) \
  {WC, CPT, NTMAX, GLO, GHI, GUARD, (const void *)&grn_field::field_kernel<WC, CPT, NTMAX, GLO, GHI, GUARD>, 0}
