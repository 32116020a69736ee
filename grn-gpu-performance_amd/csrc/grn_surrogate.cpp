// grn_surrogate.cpp -- host side of the learned pre-screen: model encoding,
// the device-resident sample store, the training loop (gather, forward,
// backward, Adam, all on the device) and the C ABI.
#include "grn_surrogate.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <thread>

#include "grn_pool.hpp"
#include "grn_search.hpp"

namespace grn {

#define SUR_TRY(expr)                                                                                 \
  do {                                                                                                \
    hipError_t e_ = (expr);                                                                           \
    if (e_ != hipSuccess)                                                                             \
      return set_last_error(GRN_E_DEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                            __LINE__);                                                                \
  } while (0)

uint16_t float_to_bf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

void encode_model(const Model &m, int n_morphogens, float *out) {
  std::fill(out, out + MLP_WIDTH, 0.f);
  const std::vector<int> in_use = m.labels_in_use(n_morphogens);
  const int ns = std::min((int)in_use.size(), MLP_GENES);
  auto slot_of = [&](int label) {
    for (int s = 0; s < ns; ++s)
      if (in_use[s] == label) return s;
    return -1;
  };
  for (int s = 0; s < ns; ++s) {
    const Gene *g = m.gene_with_label(in_use[s]);
    if (!g) continue;  // a regulator label with no product: no parameters to encode
    out[3 * s] = (float)g->lim;
    out[3 * s + 1] = (float)g->deg;
    out[3 * s + 2] = (float)(10.0 * g->dif);
  }
  for (const Link &l : m.links) {
    const int r = slot_of(l.regulator), t = slot_of(l.regulated);
    if (r < 0 || t < 0) continue;
    float *f = out + 3 * MLP_GENES + 3 * (MLP_GENES * r + t);
    f[0] = l.hill_coef < 0 ? -1.f : (l.is_and ? 0.5f : 1.f);
    f[1] = (float)l.dis_const;
    f[2] = (float)(std::fabs(l.hill_coef) / 4.0);
  }
}

template <class T>
int dev_alloc(T **p, size_t count) {
  if (hipMalloc((void **)p, sizeof(T) * std::max(count, (size_t)1)) != hipSuccess)
    return set_last_error(GRN_E_NOMEM, "hipMalloc of %zu bytes failed", sizeof(T) * count);
  return 0;
}

namespace {

// rows (n x MLP_FEATURES floats) -> bf16 rows of MLP_WIDTH, zero-padded to rows_pad
void to_bf16_rows(int n, int rows_pad, const float *features, std::vector<uint16_t> *out) {
  out->assign((size_t)rows_pad * MLP_WIDTH, 0);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < MLP_FEATURES; ++k)
      (*out)[(size_t)i * MLP_WIDTH + k] = float_to_bf16(features[(size_t)i * MLP_FEATURES + k]);
}

int round_rows(int n) { return (n + MLP_ROW_ALIGN - 1) / MLP_ROW_ALIGN * MLP_ROW_ALIGN; }

}  // namespace

Surrogate::~Surrogate() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  void *bufs[] = {d_p_, d_m_, d_v_, d_g_, d_wt_, d_w_, d_wout_, d_feat_, d_target_, d_x_, d_xt_, d_t_,
                  d_mask_, d_y_, d_dy_, d_sq_, d_loss_, d_idx_, d_ws_};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  for (int l = 0; l < MLP_HIDDEN_LAYERS; ++l) {
    void *lb[] = {d_z_[l], d_h_[l], d_ht_[l], d_dz_[l], d_dzt_[l]};
    for (void *b : lb)
      if (b) (void)hipFree(b);
  }
  if (h_rows_) (void)hipHostFree(h_rows_);
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

int Surrogate::alloc() {
  int rc;
  const size_t P = MLP_N_PARAMS, MAT = (size_t)MLP_HIDDEN_LAYERS * MLP_MAT;
  if ((rc = dev_alloc(&d_p_, P)) || (rc = dev_alloc(&d_m_, P)) || (rc = dev_alloc(&d_v_, P)) ||
      (rc = dev_alloc(&d_g_, P)) || (rc = dev_alloc(&d_wt_, MAT)) || (rc = dev_alloc(&d_w_, MAT)) ||
      (rc = dev_alloc(&d_wout_, MLP_WIDTH)))
    return rc;
  if ((rc = dev_alloc(&d_feat_, (size_t)p_.capacity * MLP_WIDTH)) || (rc = dev_alloc(&d_target_, p_.capacity)))
    return rc;
  if ((rc = dev_alloc(&d_ws_, (size_t)kMaxKSplit * MLP_MAT))) return rc;
  const size_t R = rows_cap_, RW = (size_t)rows_cap_ * MLP_WIDTH;
  if ((rc = dev_alloc(&d_x_, RW)) || (rc = dev_alloc(&d_xt_, RW)) || (rc = dev_alloc(&d_t_, R)) ||
      (rc = dev_alloc(&d_mask_, R)) || (rc = dev_alloc(&d_y_, R)) || (rc = dev_alloc(&d_dy_, R)) ||
      (rc = dev_alloc(&d_sq_, R)) || (rc = dev_alloc(&d_loss_, 1)))
    return rc;
  for (int l = 0; l < MLP_HIDDEN_LAYERS; ++l)
    if ((rc = dev_alloc(&d_z_[l], RW)) || (rc = dev_alloc(&d_h_[l], RW)) || (rc = dev_alloc(&d_ht_[l], RW)) ||
        (rc = dev_alloc(&d_dz_[l], RW)) || (rc = dev_alloc(&d_dzt_[l], RW)))
      return rc;
  return 0;
}

int Surrogate::upload_bf16_copies() {
  const size_t MAT = (size_t)MLP_HIDDEN_LAYERS * MLP_MAT;
  std::vector<uint16_t> wt(MAT), w(MAT), wout(MLP_WIDTH);
  for (int l = 0; l < MLP_HIDDEN_LAYERS; ++l)
    for (int o = 0; o < MLP_WIDTH; ++o)
      for (int i = 0; i < MLP_WIDTH; ++i) {
        const uint16_t b = float_to_bf16(host_p_[(size_t)l * MLP_MAT + (size_t)o * MLP_WIDTH + i]);
        wt[(size_t)l * MLP_MAT + (size_t)o * MLP_WIDTH + i] = b;
        w[(size_t)l * MLP_MAT + (size_t)i * MLP_WIDTH + o] = b;
      }
  for (int k = 0; k < MLP_WIDTH; ++k) wout[k] = float_to_bf16(host_p_[MLP_OFF_OUT_W + k]);
  SUR_TRY(hipMemcpyAsync(d_p_, host_p_.data(), sizeof(float) * MLP_N_PARAMS, hipMemcpyHostToDevice, stream_));
  SUR_TRY(hipMemcpyAsync(d_wt_, wt.data(), sizeof(uint16_t) * MAT, hipMemcpyHostToDevice, stream_));
  SUR_TRY(hipMemcpyAsync(d_w_, w.data(), sizeof(uint16_t) * MAT, hipMemcpyHostToDevice, stream_));
  SUR_TRY(hipMemcpyAsync(d_wout_, wout.data(), sizeof(uint16_t) * MLP_WIDTH, hipMemcpyHostToDevice, stream_));
  SUR_TRY(hipStreamSynchronize(stream_));
  return 0;
}

int Surrogate::init(int device, const grn_surrogate_params &p) {
  p_ = p;
  device_ = device;
  rows_cap_ = std::max(round_rows(p_.batch), 8192);
  SUR_TRY(hipSetDevice(device));
  SUR_TRY(hipStreamCreate(&stream_));
  SUR_TRY(hipEventCreate(&ev0_));
  SUR_TRY(hipEventCreate(&ev1_));
  int rc = alloc();
  if (rc) return rc;
  // He-uniform hidden weights (variance 2 / fan_in), output weights with
  // variance 1 / fan_in, zero biases.
  Rng rng(p_.seed);
  rng_ = p_.seed ^ 0x5DEECE66DULL;
  host_p_.assign(MLP_N_PARAMS, 0.f);
  const double a = std::sqrt(6.0 / MLP_WIDTH), ao = std::sqrt(3.0 / MLP_WIDTH);
  for (long i = 0; i < MLP_OFF_BIAS; ++i) host_p_[i] = (float)rng.uniform(-a, a);
  for (int k = 0; k < MLP_WIDTH; ++k) host_p_[MLP_OFF_OUT_W + k] = (float)rng.uniform(-ao, ao);
  SUR_TRY(hipMemsetAsync(d_m_, 0, sizeof(float) * MLP_N_PARAMS, stream_));
  SUR_TRY(hipMemsetAsync(d_v_, 0, sizeof(float) * MLP_N_PARAMS, stream_));
  std::vector<float> ones(rows_cap_, 1.f);
  SUR_TRY(hipMemcpyAsync(d_mask_, ones.data(), sizeof(float) * rows_cap_, hipMemcpyHostToDevice, stream_));
  stats_.mu = mu_;
  stats_.sigma = sigma_;
  return upload_bf16_copies();
}

int Surrogate::forward(int rows, bool keep) {
  for (int l = 0; l < MLP_HIDDEN_LAYERS; ++l) {
    MlpGemm g{};
    g.A = l == 0 ? d_x_ : d_h_[l - 1];
    g.Bt = d_wt_ + (size_t)l * MLP_MAT;
    g.M = rows;
    g.N = MLP_WIDTH;
    g.K = MLP_WIDTH;
    g.C = keep ? d_z_[l] : nullptr;
    g.bias = d_p_ + MLP_OFF_BIAS + (size_t)l * MLP_WIDTH;
    g.H = d_h_[l];
    g.Ht = keep ? d_ht_[l] : nullptr;
    SUR_TRY(grn_mlp_gemm(EPI_FWD, g, stream_));
  }
  SUR_TRY(grn_mlp_out_forward(d_h_[MLP_HIDDEN_LAYERS - 1], d_wout_, d_p_ + MLP_OFF_OUT_B, d_y_, rows, stream_));
  return 0;
}

int Surrogate::backward(int rows) {
  // the number of valid rows is folded into the loss gradient by the caller
  const int L = MLP_HIDDEN_LAYERS;
  for (int l = L - 1; l >= 0; --l) {
    MlpGemm g{};  // dW_l^T = dZ_l^T . H_{l-1}   (K = rows)
    g.A = d_dzt_[l];
    g.Bt = l == 0 ? d_xt_ : d_ht_[l - 1];
    g.M = MLP_WIDTH;
    g.N = MLP_WIDTH;
    g.K = rows;
    g.C = d_g_ + (size_t)l * MLP_MAT;
    // 16 output tiles only: slice K so that 16 x ksplit workgroups share it
    int ks = kMaxKSplit;
    while (ks > 1 && (rows % (32 * ks) || rows / ks < 128)) ks /= 2;
    g.ksplit = ks;
    g.Cpart = d_ws_;
    SUR_TRY(grn_mlp_gemm(EPI_F32, g, stream_));
    SUR_TRY(grn_mlp_row_dot(d_dzt_[l], nullptr, MLP_WIDTH, rows, d_g_ + MLP_OFF_BIAS + (size_t)l * MLP_WIDTH,
                            stream_));
    if (l > 0) {
      MlpGemm d{};  // dZ_{l-1} = (dZ_l . W_l) * gelu'(Z_{l-1})
      d.A = d_dz_[l];
      d.Bt = d_w_ + (size_t)l * MLP_MAT;
      d.M = rows;
      d.N = MLP_WIDTH;
      d.K = MLP_WIDTH;
      d.Zprev = d_z_[l - 1];
      d.H = d_dz_[l - 1];
      d.Ht = d_dzt_[l - 1];
      SUR_TRY(grn_mlp_gemm(EPI_BWD, d, stream_));
    }
  }
  return 0;
}

int Surrogate::adam() {
  ++adam_t_;
  const double b1t = 1.0 - std::pow(p_.beta1, (double)adam_t_), b2t = 1.0 - std::pow(p_.beta2, (double)adam_t_);
  MlpAdam a{};
  a.p = d_p_;
  a.m = d_m_;
  a.v = d_v_;
  a.grad = d_g_;
  a.n = MLP_N_PARAMS;
  a.beta1 = (float)p_.beta1;
  a.beta2 = (float)p_.beta2;
  a.lr_t = (float)(p_.lr * std::sqrt(b2t) / b1t);
  a.eps_t = (float)(p_.eps * std::sqrt(b2t));
  a.wt = d_wt_;
  a.w = d_w_;
  a.w_out = d_wout_;
  a.out_w_offset = MLP_OFF_OUT_W;
  SUR_TRY(grn_mlp_adam(a, stream_));
  return 0;
}

int Surrogate::add(int n, const float *features, const double *log_errors) {
  if (n <= 0) return 0;
  std::vector<uint16_t> rows;
  to_bf16_rows(n, n, features, &rows);
  return add_rows(n, rows.data(), log_errors);
}

int Surrogate::add_rows(int n, const uint16_t *rows, const double *log_errors) {
  if (pair_w_ > 0.f) return add_pairs(n, rows, rows, log_errors, log_errors);
  if (int rc0 = finish()) return rc0;
  if (n <= 0) return 0;
  std::vector<float> t(n);
  for (int i = 0; i < n; ++i) t[i] = (float)log_errors[i];
  return store_rows(n, rows, t.data());
}

int Surrogate::set_pair_weight(double w) {
  if (!(w >= 0.0) || !std::isfinite(w)) return set_last_error(GRN_E_INVALID, "pair weight %g", w);
  if (n_samples_ > 0 && (w > 0.0) != (pair_w_ > 0.f))
    return set_last_error(GRN_E_INVALID, "the store's layout (pairs or single rows) is fixed once it holds samples");
  if (w > 0.0 && p_.capacity % 2)
    return set_last_error(GRN_E_INVALID, "pair mode needs an even store capacity (%ld)", p_.capacity);
  pair_w_ = (float)w;
  return 0;
}

int Surrogate::add_pairs(int n, const uint16_t *child_rows, const uint16_t *parent_rows,
                         const double *child_log_errors, const double *parent_log_errors) {
  if (pair_w_ <= 0.f) return set_last_error(GRN_E_INVALID, "pairs need a positive pair weight");
  if (int rc0 = finish()) return rc0;
  if (n <= 0) return 0;
  // interleaved: child 2i, parent 2i + 1 (the store and its capacity stay
  // pair-aligned, so a pair never straddles the ring's end)
  std::vector<uint16_t> rows((size_t)2 * n * MLP_WIDTH);
  std::vector<float> t((size_t)2 * n);
  for (int i = 0; i < n; ++i) {
    std::memcpy(rows.data() + (size_t)(2 * i) * MLP_WIDTH, child_rows + (size_t)i * MLP_WIDTH,
                sizeof(uint16_t) * MLP_WIDTH);
    std::memcpy(rows.data() + (size_t)(2 * i + 1) * MLP_WIDTH, parent_rows + (size_t)i * MLP_WIDTH,
                sizeof(uint16_t) * MLP_WIDTH);
    t[2 * i] = (float)child_log_errors[i];
    t[2 * i + 1] = (float)parent_log_errors[i];
  }
  return store_rows(2 * n, rows.data(), t.data());
}

int Surrogate::store_rows(int n, const uint16_t *rows, const float *t) {
  if ((long)host_target_.size() < p_.capacity) host_target_.resize(std::min<long>(p_.capacity, n_samples_ + n));
  int done = 0;
  while (done < n) {
    const int k = (int)std::min<long>(n - done, p_.capacity - next_slot_);
    SUR_TRY(hipMemcpyAsync(d_feat_ + (size_t)next_slot_ * MLP_WIDTH, rows + (size_t)done * MLP_WIDTH,
                           sizeof(uint16_t) * (size_t)k * MLP_WIDTH, hipMemcpyHostToDevice, stream_));
    SUR_TRY(hipMemcpyAsync(d_target_ + next_slot_, t + done, sizeof(float) * k, hipMemcpyHostToDevice, stream_));
    for (int i = 0; i < k; ++i) host_target_[next_slot_ + i] = t[done + i];
    next_slot_ = (next_slot_ + k) % p_.capacity;
    done += k;
  }
  n_samples_ = std::min<long>(p_.capacity, n_samples_ + n);
  SUR_TRY(hipStreamSynchronize(stream_));  // the staging buffers belong to the caller
  stats_.samples = n_samples_;
  return 0;
}

void encode_model_bf16(const Model &m, int n_morphogens, uint16_t *row) {
  float f[MLP_WIDTH];
  encode_model(m, n_morphogens, f);
  for (int k = 0; k < MLP_WIDTH; ++k) row[k] = float_to_bf16(f[k]);
}

void Surrogate::encode_rows(const std::vector<const Model *> &models, std::vector<uint16_t> *rows) const {
  rows->assign(models.size() * MLP_WIDTH, 0);
  encode_rows(models, rows->data());
}

void Surrogate::encode_rows(const std::vector<const Model *> &models, uint16_t *rows) const {
  const size_t n = models.size();
  // (the host pool's threads, at least 256 rows per chunk)
  const int n_chunks = (int)std::max<size_t>(1, std::min<size_t>(host_threads(), (n + 255) / 256));
  parallel_for(n_chunks, [&](int t) {
    for (size_t i = t; i < n; i += n_chunks) encode_model_bf16(*models[i], p_.n_morphogens, rows + i * MLP_WIDTH);
  });
}

uint16_t *Surrogate::staging_rows(size_t n) {
  if (n > h_rows_cap_) {
    if (h_rows_) (void)hipHostFree(h_rows_);
    h_rows_ = nullptr;
    h_rows_cap_ = 0;
    if (hipHostMalloc((void **)&h_rows_, sizeof(uint16_t) * n * MLP_WIDTH, hipHostMallocDefault) != hipSuccess) {
      h_rows_ = nullptr;
      set_last_error(GRN_E_NOMEM, "hipHostMalloc of %zu staging rows failed", n);
      return nullptr;
    }
    h_rows_cap_ = n;
  }
  return h_rows_;
}

int Surrogate::train(int steps, double *last_loss) {
  int rc = train_async(steps);
  if (!rc) rc = finish();
  if (!rc && last_loss) *last_loss = stats_.last_loss;
  return rc;
}

int Surrogate::train_async(int steps) {
  int rc = finish();
  if (rc || steps <= 0) return rc;
  if (n_samples_ < 1) return set_last_error(GRN_E_INVALID, "the surrogate has no samples to train on");
  // standardise the targets over the store
  double s = 0, s2 = 0;
  for (long i = 0; i < n_samples_; ++i) {
    s += host_target_[i];
    s2 += (double)host_target_[i] * host_target_[i];
  }
  const double mean = s / n_samples_, var = std::max(s2 / n_samples_ - mean * mean, 0.0);
  mu_ = (float)mean;
  sigma_ = (float)std::max(std::sqrt(var), 1e-3);
  const int B = p_.batch;
  host_idx_.resize((size_t)steps * B);
  Rng rng(rng_);
  if (pair_w_ > 0.f) {  // whole pairs: rows 2j and 2j + 1 of the store
    const uint64_t n_pairs = (uint64_t)std::max<long>(n_samples_ / 2, 1);
    for (size_t b = 0; b + 1 < host_idx_.size(); b += 2) {
      const int j = (int)(rng.next() % n_pairs);
      host_idx_[b] = 2 * j;
      host_idx_[b + 1] = 2 * j + 1;
    }
  } else {
    for (int &v : host_idx_) v = (int)(rng.next() % (uint64_t)n_samples_);
  }
  rng_ = rng.next();
  if (host_idx_.size() > idx_cap_) {
    if (d_idx_) (void)hipFree(d_idx_);
    d_idx_ = nullptr;
    idx_cap_ = 0;
    if ((rc = dev_alloc(&d_idx_, host_idx_.size()))) return rc;
    idx_cap_ = host_idx_.size();
  }
  SUR_TRY(hipMemcpyAsync(d_idx_, host_idx_.data(), sizeof(int) * host_idx_.size(), hipMemcpyHostToDevice, stream_));
  SUR_TRY(hipEventRecord(ev0_, stream_));
  training_ = true;  // from here on finish() must wait, also on the error paths
  for (int st = 0; st < steps; ++st) {
    SUR_TRY(grn_mlp_gather(d_feat_, d_target_, d_idx_ + (size_t)st * B, B, mu_, 1.f / sigma_, d_x_, d_xt_, d_t_,
                           stream_));
    if ((rc = forward(B, true))) return rc;
    SUR_TRY(grn_mlp_out_backward(d_y_, d_t_, d_mask_, d_wout_, d_z_[MLP_HIDDEN_LAYERS - 1], 1.f / B, pair_w_, B, d_dy_,
                                 d_sq_,
                                 d_dz_[MLP_HIDDEN_LAYERS - 1], d_dzt_[MLP_HIDDEN_LAYERS - 1], stream_));
    SUR_TRY(grn_mlp_row_dot(d_ht_[MLP_HIDDEN_LAYERS - 1], d_dy_, MLP_WIDTH, B, d_g_ + MLP_OFF_OUT_W, stream_));
    SUR_TRY(grn_mlp_sum(d_dy_, B, d_g_ + MLP_OFF_OUT_B, stream_));
    if ((rc = backward(B))) return rc;
    if ((rc = adam())) return rc;
  }
  SUR_TRY(grn_mlp_sum(d_sq_, B, d_loss_, stream_));
  SUR_TRY(hipEventRecord(ev1_, stream_));
  return 0;
}

int Surrogate::finish() {
  if (!training_) return 0;
  training_ = false;
  float loss = 0.f, ms = 0.f;
  SUR_TRY(hipMemcpyAsync(&loss, d_loss_, sizeof(float), hipMemcpyDeviceToHost, stream_));
  SUR_TRY(hipStreamSynchronize(stream_));
  SUR_TRY(hipEventElapsedTime(&ms, ev0_, ev1_));
  stats_.last_loss = loss / p_.batch;
  stats_.train_steps = adam_t_;
  stats_.train_ms = ms;
  stats_.mu = mu_;
  stats_.sigma = sigma_;
  return 0;
}

int Surrogate::predict(int n, const float *features, double *log_errors) {
  std::vector<uint16_t> rows;
  to_bf16_rows(n, n, features, &rows);
  return predict_rows(n, rows.data(), log_errors);
}

int Surrogate::predict_rows(int n, const uint16_t *rows, double *log_errors) {
  if (int rc0 = finish()) return rc0;
  SUR_TRY(hipEventRecord(ev0_, stream_));
  // chunks of rows_cap_ rows follow one another on the stream (each upload
  // waits for the previous chunk's forward pass); one wait at the end
  std::vector<float> &y = host_y_;  // a member: outlives an early return with copies in flight
  y.resize(std::max(n, 1));
  for (int b = 0; b < n; b += rows_cap_) {
    const int k = std::min(rows_cap_, n - b), kp = round_rows(k);
    SUR_TRY(hipMemcpyAsync(d_x_, rows + (size_t)b * MLP_WIDTH, sizeof(uint16_t) * (size_t)k * MLP_WIDTH,
                           hipMemcpyHostToDevice, stream_));
    if (kp > k)  // padding rows: zero features
      SUR_TRY(hipMemsetAsync(d_x_ + (size_t)k * MLP_WIDTH, 0, sizeof(uint16_t) * (size_t)(kp - k) * MLP_WIDTH,
                             stream_));
    int rc = forward(kp, false);
    if (rc) {
      (void)hipStreamSynchronize(stream_);
      return rc;
    }
    SUR_TRY(hipMemcpyAsync(y.data() + b, d_y_, sizeof(float) * k, hipMemcpyDeviceToHost, stream_));
  }
  SUR_TRY(hipEventRecord(ev1_, stream_));
  SUR_TRY(hipStreamSynchronize(stream_));
  for (int i = 0; i < n; ++i) log_errors[i] = (double)y[i] * sigma_ + mu_;
  float ms = 0.f;
  SUR_TRY(hipEventElapsedTime(&ms, ev0_, ev1_));
  stats_.predict_ms = ms;
  return 0;
}

int Surrogate::gradient(int n, const float *features, const double *targets, float *grad, double *loss) {
  if (int rc0 = finish()) return rc0;
  if (n < 1 || round_rows(n) > rows_cap_)
    return set_last_error(GRN_E_INVALID, "gradient takes 1 to %d rows", rows_cap_);
  if (pair_w_ > 0.f && n % 2) return set_last_error(GRN_E_INVALID, "pair mode: rows come in pairs (%d rows)", n);
  const int rows = round_rows(n);
  std::vector<uint16_t> x, xt((size_t)rows * MLP_WIDTH);
  to_bf16_rows(n, rows, features, &x);
  for (int r = 0; r < rows; ++r)
    for (int k = 0; k < MLP_WIDTH; ++k) xt[(size_t)k * rows + r] = x[(size_t)r * MLP_WIDTH + k];
  std::vector<float> t(rows, 0.f), mask(rows, 0.f);
  for (int i = 0; i < n; ++i) {
    t[i] = (float)targets[i];
    mask[i] = 1.f;
  }
  SUR_TRY(hipMemcpyAsync(d_x_, x.data(), sizeof(uint16_t) * x.size(), hipMemcpyHostToDevice, stream_));
  SUR_TRY(hipMemcpyAsync(d_xt_, xt.data(), sizeof(uint16_t) * xt.size(), hipMemcpyHostToDevice, stream_));
  SUR_TRY(hipMemcpyAsync(d_t_, t.data(), sizeof(float) * rows, hipMemcpyHostToDevice, stream_));
  SUR_TRY(hipMemcpyAsync(d_mask_, mask.data(), sizeof(float) * rows, hipMemcpyHostToDevice, stream_));
  int rc = forward(rows, true);
  if (!rc) {
    SUR_TRY(grn_mlp_out_backward(d_y_, d_t_, d_mask_, d_wout_, d_z_[MLP_HIDDEN_LAYERS - 1], 1.f / n, pair_w_, rows, d_dy_,
                                 d_sq_, d_dz_[MLP_HIDDEN_LAYERS - 1], d_dzt_[MLP_HIDDEN_LAYERS - 1], stream_));
    SUR_TRY(grn_mlp_row_dot(d_ht_[MLP_HIDDEN_LAYERS - 1], d_dy_, MLP_WIDTH, rows, d_g_ + MLP_OFF_OUT_W, stream_));
    SUR_TRY(grn_mlp_sum(d_dy_, rows, d_g_ + MLP_OFF_OUT_B, stream_));
    rc = backward(rows);
  }
  if (!rc) SUR_TRY(grn_mlp_sum(d_sq_, rows, d_loss_, stream_));
  float l = 0.f;
  if (!rc) {
    SUR_TRY(hipMemcpyAsync(grad, d_g_, sizeof(float) * MLP_N_PARAMS, hipMemcpyDeviceToHost, stream_));
    SUR_TRY(hipMemcpyAsync(&l, d_loss_, sizeof(float), hipMemcpyDeviceToHost, stream_));
  }
  // restore the all-ones training mask
  std::vector<float> ones(rows_cap_, 1.f);
  SUR_TRY(hipMemcpyAsync(d_mask_, ones.data(), sizeof(float) * rows_cap_, hipMemcpyHostToDevice, stream_));
  SUR_TRY(hipStreamSynchronize(stream_));
  if (rc) return rc;
  if (loss) *loss = (double)l / n;
  return 0;
}

int Surrogate::get_params(float *p) const {
  SUR_TRY(hipStreamSynchronize(stream_));  // pending training steps included
  SUR_TRY(hipMemcpyAsync(p, d_p_, sizeof(float) * MLP_N_PARAMS, hipMemcpyDeviceToHost, stream_));
  SUR_TRY(hipStreamSynchronize(stream_));
  return 0;
}

int Surrogate::set_params(const float *p) {
  if (int rc0 = finish()) return rc0;
  host_p_.assign(p, p + MLP_N_PARAMS);
  return upload_bf16_copies();
}

}  // namespace grn

// ---- C ABI ------------------------------------------------------------------

namespace {

grn::Surrogate *as_sur(void *p) { return static_cast<grn::Surrogate *>(p); }

int encode_models(int n, const grn_model *models, int n_morph, std::vector<float> *out) {
  out->assign((size_t)n * MLP_FEATURES, 0.f);
  float row[MLP_WIDTH];
  for (int i = 0; i < n; ++i) {
    grn::Model m;
    if (!grn::model_from_abi(&models[i], &m)) return grn::set_last_error(GRN_E_INVALID, "invalid grn_model %d", i);
    grn::encode_model(m, n_morph, row);
    std::memcpy(out->data() + (size_t)i * MLP_FEATURES, row, sizeof(float) * MLP_FEATURES);
  }
  return 0;
}

}  // namespace

extern "C" {

void grn_surrogate_default_params(grn_surrogate_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof *p);
  p->n_morphogens = 3;
  p->batch = 4096;
  p->capacity = 1 << 20;
  p->lr = 1e-3;
  p->beta1 = 0.9;
  p->beta2 = 0.999;
  p->eps = 1e-8;
  p->seed = 1;
}

long grn_surrogate_n_params(void) { return MLP_N_PARAMS; }

int grn_surrogate_create(int device, const grn_surrogate_params *p, void **out) {
  if (!p || !out) return grn::set_last_error(GRN_E_INVALID, "null argument");
  *out = nullptr;
  if (p->batch < MLP_ROW_ALIGN || p->batch % MLP_ROW_ALIGN || p->capacity < 1 || p->n_morphogens < 1 ||
      p->n_morphogens > MLP_GENES || !(p->lr > 0) || !(p->beta1 >= 0 && p->beta1 < 1) ||
      !(p->beta2 >= 0 && p->beta2 < 1) || !(p->eps > 0))
    return grn::set_last_error(GRN_E_INVALID,
                               "surrogate parameters: batch a positive multiple of %d, capacity >= 1, "
                               "1 <= n_morphogens <= %d, lr > 0, 0 <= beta < 1, eps > 0",
                               MLP_ROW_ALIGN, MLP_GENES);
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
    return grn::set_last_error(GRN_E_INVALID, "device %d of %d", device, count);
  std::unique_ptr<grn::Surrogate> s(new grn::Surrogate);
  int rc = s->init(device, *p);
  if (rc) return rc;
  *out = s.release();
  return 0;
}

void grn_surrogate_destroy(void *sur) { delete as_sur(sur); }

int grn_surrogate_encode(const grn_model *model, int n_morphogens, float *features) {
  if (!model || !features) return grn::set_last_error(GRN_E_INVALID, "null argument");
  std::vector<float> f;
  int rc = encode_models(1, model, n_morphogens, &f);
  if (rc) return rc;
  std::memcpy(features, f.data(), sizeof(float) * MLP_FEATURES);
  return 0;
}

int grn_surrogate_add(void *sur, int n, const grn_model *models, const double *errors) {
  grn::Surrogate *s = as_sur(sur);
  if (!s || n < 0 || (n && (!models || !errors))) return grn::set_last_error(GRN_E_INVALID, "null argument");
  std::vector<grn_model> keep;
  std::vector<double> logs;
  for (int i = 0; i < n; ++i)
    if (errors[i] > 0 && std::isfinite(errors[i])) {
      keep.push_back(models[i]);
      logs.push_back(std::log(errors[i]));
    }
  std::vector<float> f;
  int rc = encode_models((int)keep.size(), keep.data(), s->n_morphogens(), &f);
  if (rc) return rc;
  return s->add((int)keep.size(), f.data(), logs.data());
}

int grn_surrogate_train(void *sur, int steps, double *loss) {
  grn::Surrogate *s = as_sur(sur);
  if (!s) return grn::set_last_error(GRN_E_INVALID, "null surrogate");
  return s->train(steps, loss);
}

int grn_surrogate_set_pair_weight(void *sur, double weight) {
  grn::Surrogate *s = as_sur(sur);
  if (!s) return grn::set_last_error(GRN_E_INVALID, "null surrogate");
  return s->set_pair_weight(weight);
}

int grn_surrogate_predict(void *sur, int n, const grn_model *models, double *log_errors) {
  grn::Surrogate *s = as_sur(sur);
  if (!s || n < 0 || (n && (!models || !log_errors))) return grn::set_last_error(GRN_E_INVALID, "null argument");
  std::vector<float> f;
  int rc = encode_models(n, models, s->n_morphogens(), &f);
  if (rc) return rc;
  return s->predict(n, f.data(), log_errors);
}

int grn_surrogate_predict_features(void *sur, int n, const float *features, double *log_errors) {
  grn::Surrogate *s = as_sur(sur);
  if (!s || n < 0 || (n && (!features || !log_errors))) return grn::set_last_error(GRN_E_INVALID, "null argument");
  return s->predict(n, features, log_errors);
}

int grn_surrogate_gradient(void *sur, int n, const float *features, const double *targets, float *grad,
                           double *loss) {
  grn::Surrogate *s = as_sur(sur);
  if (!s || !features || !targets || !grad) return grn::set_last_error(GRN_E_INVALID, "null argument");
  return s->gradient(n, features, targets, grad, loss);
}

int grn_surrogate_get_params(void *sur, float *params) {
  grn::Surrogate *s = as_sur(sur);
  if (!s || !params) return grn::set_last_error(GRN_E_INVALID, "null argument");
  return s->get_params(params);
}

int grn_surrogate_set_params(void *sur, const float *params) {
  grn::Surrogate *s = as_sur(sur);
  if (!s || !params) return grn::set_last_error(GRN_E_INVALID, "null argument");
  return s->set_params(params);
}

int grn_surrogate_get_stats(void *sur, grn_surrogate_stats *out) {
  grn::Surrogate *s = as_sur(sur);
  if (!s || !out) return grn::set_last_error(GRN_E_INVALID, "null argument");
  int rc = s->finish();
  if (rc) return rc;
  *out = s->stats();
  return 0;
}

int grn_mlp_gemm_bf16(int device, int M, int N, int K, const unsigned short *A, const unsigned short *Bt, float *C) {
  if (!A || !Bt || !C) return grn::set_last_error(GRN_E_INVALID, "null argument");
  if (M < 128 || N < 128 || K < 32 || M % 128 || N % 128 || K % 32)
    return grn::set_last_error(GRN_E_INVALID, "M, N must be positive multiples of 128 and K of 32");
  if (hipSetDevice(device) != hipSuccess) return grn::set_last_error(GRN_E_INVALID, "device %d", device);
  __bf16 *da = nullptr, *db = nullptr;
  float *dc = nullptr;
  int rc = 0;
  if ((rc = grn::dev_alloc(&da, (size_t)M * K)) || (rc = grn::dev_alloc(&db, (size_t)N * K)) ||
      (rc = grn::dev_alloc(&dc, (size_t)M * N))) {
  } else {
    MlpGemm g{};
    g.A = da;
    g.Bt = db;
    g.M = M;
    g.N = N;
    g.K = K;
    g.C = dc;
    if (hipMemcpy(da, A, sizeof(uint16_t) * (size_t)M * K, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(db, Bt, sizeof(uint16_t) * (size_t)N * K, hipMemcpyHostToDevice) != hipSuccess ||
        grn_mlp_gemm(EPI_F32, g, 0) != hipSuccess ||
        hipMemcpy(C, dc, sizeof(float) * (size_t)M * N, hipMemcpyDeviceToHost) != hipSuccess)
      rc = grn::set_last_error(GRN_E_DEVICE, "GEMM test failed: %s", hipGetErrorString(hipGetLastError()));
  }
  if (da) (void)hipFree(da);
  if (db) (void)hipFree(db);
  if (dc) (void)hipFree(dc);
  return rc;
}

}  // extern "C"
