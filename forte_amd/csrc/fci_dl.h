// fci_dl.h -- Davidson-Liu eigensolver with device-resident vectors
// (forte/helpers/davidson_liu_solver.h:52-240, davidson_liu_solver.cc:45-731).
#pragma once

#include "fci_internal.h"

#include <functional>

namespace fci {

using SparseVec = std::vector<std::pair<int64_t, double>>;

// The guess construction of initial_guess_dets from already selected determinant addresses
// (host only: spin completion, S^2 and H blocks, per-multiplicity eigenvectors).
std::pair<std::vector<SparseVec>, std::vector<SparseVec>> initial_guess_from_selection(
    const std::vector<int64_t>& idx, int64_t nbs, int n, int na, int nb, const HostInts& hi, int multiplicity,
    int num_guess_states);

// Symmetric eigensolver (cyclic Jacobi), ascending; V column-major: eigenvector k at V[k*n..].
void jacobi_eigh(int n, const std::vector<double>& A, std::vector<double>& evals, std::vector<double>& V);

// FCISolver's determinant guess (fci_solver_initial_guess.cc:88-101, sparse_initial_guess.cc:71-264)
// from a device H diagonal: {guess vectors, other-multiplicity roots to project out}.
// d_hdiag: the diagonal used for the selection (determinants that may not be chosen hold +inf);
// nsel: how many determinants may be chosen
std::pair<std::vector<SparseVec>, std::vector<SparseVec>> initial_guess_dets(const double* d_hdiag, int64_t ndet,
                                                                             int64_t nsel, int64_t nbs, int n, int na,
                                                                             int nb, const HostInts& hi,
                                                                             int multiplicity, int ndets_per_guess,
                                                                             int num_guess_states, hipStream_t st,
                                                                             SelectScratch* ws = nullptr);

class DavidsonLiu {
  public:
    // sigma(b, sigma, nvec): nvec consecutive vectors of length size (device pointers)
    using SigmaFn = std::function<void(const double* b, double* sigma)>;
    // optional replacement of form_correction_vectors: (residuals r [nroot x size], lambda) -> r
    using PrecondFn = std::function<bool(double* r, const double* lambda, int nroot, int iter)>;

    DavidsonLiu(int64_t size, int nroot, int collapse_per_root, int subspace_per_root);

    void add_h_diag_host(const double* h, int64_t n);
    void add_h_diag_device(const double* d_h, int64_t n);
    void add_guesses(const std::vector<SparseVec>& g) { guesses_ = g; }
    void add_project_out_vectors(const std::vector<SparseVec>& v);
    void add_sigma_builder(SigmaFn f) {
        sigma_ = std::move(f);
        s32_ = false;
    }
    // sigma rows kept in FP32 (a reduced-precision phase): dots, residuals and rotations read
    // them and accumulate in FP64
    using Sigma32Fn = std::function<void(const double* b, float* sigma)>;
    void add_sigma_builder32(Sigma32Fn f);
    void set_preconditioner(PrecondFn f) { precond_ = std::move(f); }
    // Olsen's correction on top of the diagonal one (not in the reference; off by default)
    void set_olsen(bool v) { olsen_ = v; }
    // determinants allowed to be nonzero (point-group sector); random vectors are zeroed elsewhere
    void set_mask(const uint8_t* d_mask) { mask_ = d_mask; }
    // called after the residual norms of an iteration are known; returning true ends the solve
    // as converged (used to end a reduced-precision phase once it stalls)
    using IterFn = std::function<bool(int iter, const std::vector<double>& resid, const std::vector<double>& lambda)>;
    void set_iteration_observer(IterFn f) { observe_ = std::move(f); }
    void set_e_convergence(double v) { e_conv_ = v; }
    void set_r_convergence(double v) { r_conv_ = v; }
    void set_maxiter(int v) { maxiter_ = v; }
    void set_print_level(int v) { print_ = v; }
    void reset();
    bool solve();

    int64_t size() const { return size_; }
    int nroot() const { return nroot_; }
    const std::vector<double>& eigenvalues() const { return lambda_; }
    const double* eigenvector_device(int k) const { return b_.p + size_t(k) * size_; }
    void eigenvector_host(int k, double* out) const;
    const double* hdiag_device() const { return hdiag_.p; }
    int iterations() const { return iters_; }
    int sigma_builds() const { return nsigma_; }
    const std::vector<double>& residual_norms() const { return resid_; }
    // Ritz vector of root 0 of the current iteration (valid inside the preconditioner callback)
    void ritz_vector0(double* out) { rotate(b_.p, out, 1); }
    int basis_size() const { return basis_size_; }

  private:
    int64_t size_;
    int nroot_, collapse_size_, subspace_size_;
    int basis_size_ = 0, sigma_size_ = 0;
    int maxiter_ = 50, print_ = 0;
    double e_conv_ = 1.0e-12, r_conv_ = 1.0e-6;
    double discard_ = 1.0e-7, orth_ = 1.0e-12;
    SigmaFn sigma_;
    Sigma32Fn sigma32_;
    bool s32_ = false;
    DBuf<float> sigma32_v_;
    PrecondFn precond_;
    IterFn observe_;
    bool olsen_ = false;
    DBuf<double> olsen_x_;
    void olsen_correct(int k);
    const uint8_t* mask_ = nullptr;
    bool have_hdiag_ = false;
    DBuf<double> hdiag_, b_, sigma_v_, r_, temp_;
    DBuf<double> part_;  // reduction partials
    DBuf<double> small_; // coefficients uploaded per call
    DBuf<int64_t> pidx_;
    DBuf<int64_t> poff_dev_;  // device copy of poff_ (bad-root offsets)
    DBuf<double> ak_;         // Ritz coefficients of one root
    DBuf<double> pval_;
    std::vector<int64_t> poff_;
    std::vector<SparseVec> guesses_;
    std::vector<double> lambda_, lambda_old_, resid_, alpha_;
    std::vector<double> G_, S_;  // cached b.sigma and b.b (subspace x subspace)
    int iters_ = 0, nsigma_ = 0;
    int gvalid_ = 0;  // G_ entries [i][j] with i,j < gvalid_ are current
    uint64_t rng_ = 0x5EEDULL;

    // device building blocks
    void dots(const double* x, const double* Y, int m, double* out);  // out[i] = x . Y_i
    template <class TX, class TY> void dots_t(const TX* x, const TY* Y, int m, double* out);
    template <class T> void lincomb_t(const T* X, int m, const double* coef, int k, T* out);
    template <class T> void rotate_t(const T* X, T* out, int kout);
    void lincomb(const double* X, int m, const double* coef /*m x k, col-major per output*/, int k,
                 double* out);
    double add_row_and_orthonormalize(double* A, int rowsA, const double* v_src, bool* ok);
    // a0, d0 (optional): scale of vsrc and its overlaps d0_i = a0 vsrc . A_i, when already known
    double add_row_fused(double* A, int rowsA, const double* vsrc, bool* ok, double a0 = 0.0,
                         const std::vector<double>* d0 = nullptr);
    double scale_norm(double* v, double a);  // v <- a v; returns ||v||^2
    // residual of root k into r (diagonal preconditioner when precond); returns ||r||^2 before
    // preconditioning; with ov: ov[0] = ||r||^2 after it and ov[1 + i] = r . b_i (i < bs)
    double residual(int k, bool precond, std::vector<double>* ov);
    double gs_fused(double* v, const double* src, const double* A, int m, const std::vector<double>& d, double a,
                    std::vector<double>& ov);
    int add_rows_and_orthonormalize(double* A, int rowsA, const double* B, int rowsB);
    void project_out(double* v, int nvec);
    void fill_random(double* v, int nvec);
    void rotate(const double* X, double* out, int kout);  // out_k = sum_j alpha[j][k] X_j
    void form_G();
    void check_orthonormality();
    void get_results();
};

}  // namespace fci
