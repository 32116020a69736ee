// grn_tiled.hpp -- halo-tiled kernel for grids whose field does not fit the
// LDS of one workgroup (e.g. 96x96, 100x100, 128x128 with many genes).
//
// One workgroup still owns one model at a time (persistent, work queue).  The
// model's state lives in device memory as two sets of per-cell records
// (hi[0..G), lo[0..G), padding: grn_cell_rec), A (time s) and B (time s + T),
// so a window cell's whole state moves with a few 16-byte loads.  The grid is covered by tiles of tile_sx x tile_sy
// cells; each tile is processed in an extended window of tile_w x tile_h cells
// (a halo of tile_halo = T cells on every side):
//   load the window from A into registers and the LDS mirror, run T Euler
//   steps with the on-chip step code of grn_field.hpp, write the interior
//   back to B.
// A window cell at distance d from the window edge is exact for d steps, so
// after T steps the interior (d >= T) is exact; the halo is recomputed by the
// neighbouring tiles (temporal blocking: one device-memory round trip per T
// steps instead of per step).  Cells outside the grid are held at zero with a
// zero time step and the no-flux neighbour count comes from global
// coordinates.  With an early-stop threshold (minConcChange >= 0) every step
// needs the grid-wide max change, so T drops to one step per round trip.
#pragma once
#include "grn_field.hpp"

namespace grn_field {

typedef float float4v __attribute__((ext_vector_type(4)));

// Fitness terms straight from the device-memory state (c = hi + lo) of the
// target genes n_in .. n_in + count / N - 1, in [target][cell] order.
__device__ __noinline__ double score_terms_rec(const float *A, int rec, int G, int n_in, int N,
                                               const double *target, int count, double thr, int mse) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < count; i += blockDim.x) {
    const int tg = i / N, cell = i - tg * N;
    const float *r = A + (size_t)cell * rec + n_in + tg;
    const double d = ((double)r[0] + (double)r[G]) - target[i];
    if (mse) {
      acc += d * d;
    } else {
      const double u = __builtin_fabs(d) - thr;
      if (u > 0) acc += log(1.0 + u);
    }
  }
  return acc;
}

// WP > 0: the LDS window rows have the compile-time pitch WP + 2 (the window
// is at most WP cells wide), so every neighbour read has an immediate offset.
// FAST: the program's state provably stays >= 0 (GrnProgram::fast), so the
// step drops the minConc clamp and the log2 floor; the kernel switches once
// per model between the two bodies.
template <int G, int CPT, int NTMAX, bool GUARD, int WP, bool FAST>
__device__ __forceinline__ void run_tiled_model(const GrnLaunch &a, const GrnProgram *__restrict__ progs,
                                                const GrnDataset *__restrict__ datasets, int m, int pi,
                                                Shared sh) {
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
  const GrnProgram *P = progs + pi;
  const GrnDataset *D = datasets + P->dataset;
  const int W = D->W, H = D->H, N = W * H;
  const int TW = a.tile_w, TH = a.tile_h, T = a.tile_halo, SX = a.tile_sx, SY = a.tile_sy;
  const int ntx = (W + SX - 1) / SX, nty = (H + SY - 1) / SY;
  constexpr int R = grn_cell_rec(G);  // floats per cell record
  float *A = a.scratch + (size_t)blockIdx.x * a.slot_floats;  // state at time s: N records
  float *B = A + (size_t)R * N;                               // state at time s + T

  Model<WP, G, CPT, NTMAX, GUARD, true> f;
  f.P = P;
  f.sh = sh;
  f.W = TW;
  f.H = TH;
  f.plane = f.pc() * (TH + 2);
  f.dt = P->dt;
  f.min_conc = D->min_conc;
  f.clamp_any_r = !(D->min_conc > 0.f || !D->nonneg);
  const int n_active = TW * (TH / CPT);  // the launcher guarantees tile_h % CPT == 0
  f.active = tid < n_active;
  f.x = f.active ? tid % TW : 0;
  f.y0 = f.active ? (tid / TW) * CPT : 0;
  f.ubase = f.y0 * f.pc() + f.x;
  const int n_exp = D->n_exp, n_in = D->n_in, n_tgt = D->n_tgt;
  const int S = (a.step_cap > 0 && a.step_cap < D->steps) ? a.step_cap : D->steps;
  const double mcc = D->min_conc_change_d;
  const bool early = mcc >= 0.0;
  const double max_err = a.max_err[pi];
  const int plane_floats = G * f.plane;

  // The LDS window planes start at zero once per model: the one-cell ring
  // around the window (the no-flux zeros) is never written, the window cells
  // of diffusing genes are rewritten by every tile before they are read, and
  // the planes of genes that do not diffuse stay zero (their diffusion term
  // is multiplied by dif = 0, so they must hold finite values).
  for (int i = tid; i < plane_floats; i += blockDim.x) sh.lds[i] = 0.f;
  __syncthreads();  // zeroed before any tile stores its planes

  double err = 0.0, chg_acc = 0.0;
  int e = 0, steps_total = 0;
  while (e < n_exp && err <= max_err) {
    // Initial state at time 0 in A: inputs for genes < nIn, zero elsewhere.
    for (int cell = tid; cell < N; cell += blockDim.x) {
      float r[R];
#pragma unroll
      for (int j = 0; j < R; ++j) r[j] = 0.f;
#pragma unroll
      for (int g = 0; g < G; ++g)
        if (g < n_in) {
          const double v = D->init[((size_t)e * n_in + g) * N + cell];
          r[g] = (float)v;
          r[G + g] = (float)(v - (double)r[g]);
        }
      float4v *dst = reinterpret_cast<float4v *>(A + (size_t)cell * R);
#pragma unroll
      for (int j = 0; j < R / 4; ++j) dst[j] = float4v{r[4 * j], r[4 * j + 1], r[4 * j + 2], r[4 * j + 3]};
    }
    __syncthreads();

    int s = 0;
    float chg = 1.0f;
    while (s < S && (!early || (double)chg > mcc)) {
      const int Tb = early ? 1 : (S - s < T ? S - s : T);
      const bool track = early || s + Tb == S;
      float cblk = 0.f;
      for (int ty = 0; ty < nty; ++ty) {
        for (int tx = 0; tx < ntx; ++tx) {
          const int ox = tx * SX - T, oy = ty * SY - T;
          const int gx = ox + f.x;
          // cell k of the strip: global index, or -1 outside the grid
          auto gidx = [&](int k) {
            const int gy = oy + f.y0 + k;
            return (f.active && gx >= 0 && gx < W && gy >= 0 && gy < H) ? gy * W + gx : -1;
          };
          f.nfb = 0u;
          f.cellb = 0u;
#pragma unroll
          for (int k = 0; k < CPT; ++k) {
            const int yt = f.y0 + k, gy = oy + yt;
            const bool in = gidx(k) >= 0;
            const bool interior = in && f.x >= T && f.x < T + SX && yt >= T && yt < T + SY;
            const unsigned nfk = in ? (unsigned)((gx > 0) + (gx < W - 1) + (gy > 0) + (gy < H - 1)) : 0u;
            f.nfb |= nfk << (3 * k);
            f.cellb |= (in ? 1u : 0u) << k;
            f.cellb |= (interior ? 1u : 0u) << (16 + k);
          }
#pragma unroll
          for (int k = 0; k < CPT; ++k) {
            // load from an in-range record, then select: no out-of-range
            // address is ever formed for cells outside the grid
            const int gk = gidx(k);
            const float4v *src = reinterpret_cast<const float4v *>(A + (size_t)(gk >= 0 ? gk : 0) * R);
            float r[R];
#pragma unroll
            for (int j = 0; j < R / 4; ++j) {
              const float4v q = src[j];
              r[4 * j] = q.x;
              r[4 * j + 1] = q.y;
              r[4 * j + 2] = q.z;
              r[4 * j + 3] = q.w;
            }
#pragma unroll
            for (int g = 0; g < G; ++g) {
              f.hi[g][k] = gk >= 0 ? r[g] : 0.f;
              f.lo[g][k] = gk >= 0 ? r[G + g] : 0.f;
            }
          }
          // (the previous tile's last step ended with a barrier after its
          // last plane read, so the planes can be stored at once)
          f.store_planes();
          __syncthreads();
          // (the step with the per-gene diffusion branch only: a second
          // copy of the step loop inside one body costs register spills)
          float c = 0.f;
          for (int t = 0; t < Tb; ++t) {
            if (track && t == Tb - 1)
              c = f.template step<true, FAST, true>();
            else
              f.template step<false, FAST, true>();
          }
          cblk = __builtin_fmaxf(cblk, c);
#pragma unroll
          for (int k = 0; k < CPT; ++k)
            if (f.trkv(k)) {
              float r[R];
#pragma unroll
              for (int g = 0; g < G; ++g) {
                r[g] = f.hi[g][k];
                r[G + g] = f.lo[g][k];
              }
#pragma unroll
              for (int j = 2 * G; j < R; ++j) r[j] = 0.f;
              float4v *dst = reinterpret_cast<float4v *>(B + (size_t)gidx(k) * R);
#pragma unroll
              for (int j = 0; j < R / 4; ++j) dst[j] = float4v{r[4 * j], r[4 * j + 1], r[4 * j + 2], r[4 * j + 3]};
            }
        }
      }
      float *t = A;
      A = B;
      B = t;
      __syncthreads();  // the new state is complete before the next block reads it
      s += Tb;
      if (track) chg = cblk;
    }
    steps_total += s;
    chg_acc += (double)chg;

    const bool mse = D->kind == 1;
    double acc = score_terms_rec(A, R, G, n_in, N, D->target + (size_t)e * n_tgt * N, n_tgt * N, D->thr,
                                 mse ? 1 : 0);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) sh.wsum[wave] = acc;
    __syncthreads();
    double tot = 0.0;
    for (int w = 0; w < nw; ++w) tot += sh.wsum[w];
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
      for (int i = tid; i < G * N; i += blockDim.x) {
        const int g = i / N, cell = i - g * N;
        fo[i] = (double)A[(size_t)cell * R + g] + (double)A[(size_t)cell * R + G + g];
      }
    }
    ++e;
    __syncthreads();  // wsum reads and the state reads done before reuse
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

template <int CPT, int NTMAX, bool GUARD, int WP, int G, int GHI>
__device__ __forceinline__ void dispatch_tiled(int g, int fast, const GrnLaunch &a,
                                               const GrnProgram *__restrict__ progs,
                                               const GrnDataset *__restrict__ datasets, int m, int pi,
                                               Shared sh) {
  if constexpr (G <= GHI) {
    if (g == G) {
      if (fast)
        run_tiled_model<G, CPT, NTMAX, GUARD, WP, true>(a, progs, datasets, m, pi, sh);
      else
        run_tiled_model<G, CPT, NTMAX, GUARD, WP, false>(a, progs, datasets, m, pi, sh);
    } else {
      dispatch_tiled<CPT, NTMAX, GUARD, WP, G + 1, GHI>(g, fast, a, progs, datasets, m, pi, sh);
    }
  }
}

template <int CPT, int NTMAX, int GLO, int GHI, bool GUARD, int WP>
__global__ __launch_bounds__(NTMAX) void tiled_kernel(GrnLaunch a, const GrnProgram *__restrict__ progs,
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
    const int fast = __builtin_amdgcn_readfirstlane(progs[pi].fast);
    dispatch_tiled<CPT, NTMAX, GUARD, WP, GLO, GHI>(g, fast, a, progs, datasets, m, pi, sh);
    __syncthreads();
  }
}

}  // namespace grn_field

#define GRN_TILED_ENTRY(CPT, NTMAX, GLO, GHI, GUARD, WP) \
  {-1, CPT, NTMAX, GLO, GHI, GUARD, (const void *)&grn_field::tiled_kernel<CPT, NTMAX, GLO, GHI, GUARD, WP>, WP}
