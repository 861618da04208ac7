// fci_lp.h -- learned Davidson preconditioner: BF16-MFMA inference and FP32 training on gfx950.
#pragma once

#include "fci_internal.h"

namespace fci {

constexpr int LP_NIN = 34, LP_NH = 128;
constexpr int LP_OFF_W1 = 0;
constexpr int LP_OFF_B1 = LP_OFF_W1 + LP_NH * LP_NIN;
constexpr int LP_OFF_W2 = LP_OFF_B1 + LP_NH;
constexpr int LP_OFF_B2 = LP_OFF_W2 + LP_NH * LP_NH;
constexpr int LP_OFF_W3 = LP_OFF_B2 + LP_NH;
constexpr int LP_OFF_B3 = LP_OFF_W3 + LP_NH;
constexpr int LP_NPARAM = LP_OFF_B3 + 1;

// Inference copy of the weights and the in-place preconditioner r <- r / (lambda - H) * exp(MLP(x)).
class LearnedPrecond {
  public:
    void set_weights(const float* p);
    bool ready() const { return wpk_.p != nullptr; }
    // one root: r (device, ndet) is the raw residual with norm rnorm; overwritten by the correction
    void apply(double* r, const double* hdiag, double lambda, double rnorm, const SpinTables& a,
               const SpinTables& b, hipStream_t st) const;
    const std::vector<float>& params() const { return params_; }

  private:
    std::vector<float> params_;
    DBuf<uint16_t> wpk_;  // packed BF16 A operands
    DBuf<float> wf_;      // b2, w3, b3
};

// X[i][0..33] = features of determinant idx[i] (raw residual r with norm rnorm)
void lp_features_device(const SpinTables& a, const SpinTables& b, const int64_t* d_idx, int64_t m,
                        const double* d_h, double lambda, const double* d_r, double rnorm, float* d_X,
                        hipStream_t st);

// harvest helpers: d0 = r / (lambda - h) and xs = x at the sampled determinants; e = c/cx - xs
void lp_sample_device(const int64_t* d_idx, int64_t m, const double* d_h, double lambda, const double* d_r,
                      const double* d_x, float* d_d0, double* d_xs, hipStream_t st);
void lp_target_device(const int64_t* d_idx, int64_t m, const double* d_c, double cx, const double* d_xs,
                      float* d_e, hipStream_t st);
double dot_device(const double* a, const double* b, int64_t n, hipStream_t st);

// FP32 training on groups of at most `group` samples: loss = 1 - cos^2(d, e) + reg mean(y^2).
class LPTrainer {
  public:
    explicit LPTrainer(int64_t group);
    void set_params(const float* p);  // also resets the Adam moments
    void get_params(float* p) const;
    void forward(const float* X, int64_t m, hipStream_t st);   // y = MLP(X)
    float loss_grad(const float* X, const float* d0, const float* e, int64_t m, float reg, hipStream_t st);
    void adam(float lr, hipStream_t st);
    void get_grad(float* g) const;
    void get_y(float* y, int64_t m) const;
    int64_t group() const { return G_; }

  private:
    int64_t G_;
    int t_ = 0;
    DBuf<float> p_, g_, m1_, m2_, z1_, a1_, z2_, a2_, y_, gy_, loss_;
};

}  // namespace fci
