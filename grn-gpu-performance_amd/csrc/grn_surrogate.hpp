// grn_surrogate.hpp -- the learned pre-screen of offspring: a 468-feature
// encoding of a model, a device-resident store of (encoding, log error)
// samples from the simulator, and the MLP trained on them with Adam.
// No counterpart in the reference, which simulates every child.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "grn.h"
#include "grn_mlp.h"
#include "grn_model.hpp"

namespace grn {

// The 468 features of a model, in MLP_WIDTH floats (the rest zero):
//   genes in use in ascending label order (the structural genes first), at
//   most 12; gene slot s: [3s] = lim, [3s+1] = deg, [3s+2] = 10 * dif;
//   link from slot r to slot t: base 36 + 3 (12 r + t): type (+1 activator
//   OR, +0.5 activator AND, -1 repressor, 0 none), K, |n| / 4.
void encode_model(const Model &m, int n_morphogens, float *out);

// Round to nearest even, the conversion the device uses for float -> bf16.
uint16_t float_to_bf16(float f);

// encode_model straight into one BF16 row of MLP_WIDTH values.
void encode_model_bf16(const Model &m, int n_morphogens, uint16_t *row);

class Surrogate {
 public:
  ~Surrogate();
  int init(int device, const grn_surrogate_params &p);

  int add(int n, const float *features, const double *log_errors);
  int train(int steps, double *last_loss);
  // Enqueues the training steps and returns at once; finish() (called by
  // every other member that touches the device) waits for them and records
  // the loss and time.
  int train_async(int steps);
  int finish();
  int predict(int n, const float *features, double *log_errors);
  // The same on rows already encoded as BF16 (n x MLP_WIDTH).  In pair mode
  // each row is stored as a pair with itself (no pairwise information).
  int add_rows(int n, const uint16_t *rows, const double *log_errors);
  // Pair mode (pair weight > 0): the store holds (child, parent) pairs and
  // training adds the pairwise term of grn_mlp_out_backward, which fits the
  // predicted change from parent to child to the simulated one.  The weight
  // can change while the store is empty, or between two positive values.
  int set_pair_weight(double w);
  double pair_weight() const { return pair_w_; }
  // n pairs: child rows and parent rows (n x MLP_WIDTH each) with their ln errors.
  int add_pairs(int n, const uint16_t *child_rows, const uint16_t *parent_rows, const double *child_log_errors,
                const double *parent_log_errors);
  int predict_rows(int n, const uint16_t *rows, double *log_errors);
  // Encodes models straight into BF16 rows (n x MLP_WIDTH), on host threads.
  void encode_rows(const std::vector<const Model *> &models, std::vector<uint16_t> *rows) const;
  void encode_rows(const std::vector<const Model *> &models, uint16_t *rows) const;
  // A pinned host buffer of at least n rows (MLP_WIDTH BF16 values each) for
  // feature uploads, kept until the surrogate is destroyed; n = 0 returns the
  // current buffer.  Null (with the error set) if the allocation fails.
  uint16_t *staging_rows(size_t n);
  // One forward and backward pass on given rows (standardised targets), no
  // update: loss and the gradient of every parameter.
  int gradient(int n, const float *features, const double *targets, float *grad, double *loss);
  int get_params(float *p) const;
  int set_params(const float *p);
  grn_surrogate_stats stats() const { return stats_; }
  int n_morphogens() const { return p_.n_morphogens; }

 private:
  int alloc();
  int store_rows(int n, const uint16_t *rows, const float *targets);
  int upload_bf16_copies();
  int forward(int rows, bool keep_for_backward);
  int backward(int rows);
  int adam();

  grn_surrogate_params p_{};
  int device_ = 0;
  hipStream_t stream_ = nullptr;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  std::vector<float> host_p_;  // initial parameters (seeded)
  // parameters and optimizer state (FP32), gradient
  float *d_p_ = nullptr, *d_m_ = nullptr, *d_v_ = nullptr, *d_g_ = nullptr;
  // partial weight gradients of the K-sliced GEMM
  static constexpr int kMaxKSplit = 16;
  float *d_ws_ = nullptr;
  __bf16 *d_wt_ = nullptr, *d_w_ = nullptr, *d_wout_ = nullptr;
  // sample store
  __bf16 *d_feat_ = nullptr;
  float *d_target_ = nullptr;
  std::vector<float> host_target_;  // copy of the targets, for standardisation
  long n_samples_ = 0, next_slot_ = 0;
  // activations for up to rows_cap_ rows
  int rows_cap_ = 0;
  __bf16 *d_x_ = nullptr, *d_xt_ = nullptr;
  float *d_t_ = nullptr, *d_mask_ = nullptr, *d_y_ = nullptr, *d_dy_ = nullptr, *d_sq_ = nullptr, *d_loss_ = nullptr;
  float *d_z_[MLP_HIDDEN_LAYERS] = {};
  __bf16 *d_h_[MLP_HIDDEN_LAYERS] = {}, *d_ht_[MLP_HIDDEN_LAYERS] = {};
  __bf16 *d_dz_[MLP_HIDDEN_LAYERS] = {}, *d_dzt_[MLP_HIDDEN_LAYERS] = {};
  int *d_idx_ = nullptr;
  size_t idx_cap_ = 0;
  std::vector<int> host_idx_;  // source of the pending index upload
  uint16_t *h_rows_ = nullptr;  // pinned staging rows
  std::vector<float> host_y_;   // predictions copied back
  size_t h_rows_cap_ = 0;
  bool training_ = false;      // steps enqueued, finish() not yet called
  uint64_t rng_ = 0;
  long adam_t_ = 0;
  float mu_ = 0.f, sigma_ = 1.f;
  float pair_w_ = 0.f;
  grn_surrogate_stats stats_{};
};

}  // namespace grn
