/*
 * forte_fci.h -- C ABI of the gfx950 FCI solver (libforte_fci_gfx950.so).
 *
 * The drop-in boundary is Forte's ActiveSpaceMethod / FCISolver interface
 * (forte/base_classes/active_space_method.h:95-388, forte/fci/fci_solver.h:51-246). Forte has no
 * C ABI of its own; these entry points are what a binding of that interface needs (a ctypes /
 * pybind11 / cgo stub calls them with plain pointers). Every function returns 0 on success and
 * a negative status on failure; fci_last_error() then holds the message. No C++ exception
 * crosses this boundary. Output buffers are owned by the caller.
 *
 * Scope: all orbitals active (doubly occupied ones are folded into the integrals by the caller),
 * any na / nb, C1 or one irrep of a point group (fci_set_symmetry); n <= 22 orbitals for the string
 * tables and n <= 19 for sigma (the alpha-beta kernels hold at most 12 row tiles of 16 packed pairs).
 * Integrals use Forte's
 * layout (forte/integrals/active_space_integrals.h:103-127): oei n*n row-major, tei n^4
 * row-major physicist, tei_aa / tei_bb antisymmetrized <pq||rs>, tei_ab = <pq|rs>.
 * Determinants are ordered alpha-major, C[Ia*Nb + Ib] (forte/fci/fci_vector.cc:108-114).
 */
#ifndef FORTE_FCI_H
#define FORTE_FCI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fci_solver fci_solver;
typedef struct fci_dl fci_dl;

enum fci_status {
    FCI_OK = 0,
    FCI_ERR_ARG = -1,     /* invalid argument */
    FCI_ERR_STATE = -2,   /* call out of order (e.g. rdms before compute_energy) */
    FCI_ERR_DEVICE = -3,  /* HIP runtime error or no gfx950 device */
    FCI_ERR_CONVERGE = -4,/* Davidson-Liu did not converge and die_if_not_converged is set */
    FCI_ERR_NUMERIC = -5  /* Davidson-Liu numerical failure (linear dependence, ...) */
};

enum fci_precision {
    FCI_PREC_FP64 = 0,  /* every sigma in FP64 (Forte's arithmetic) */
    FCI_PREC_MIXED = 1  /* FP32 C/sigma storage, BF16x3 MFMA contractions accumulated in FP32,
                           FP64 Davidson algebra, FP64 refinement of the converged vector */
};

enum fci_preconditioner {
    FCI_PRECOND_DIAGONAL = 0, /* r / (lambda - H_II), davidson_liu_solver.cc:432-448 */
    FCI_PRECOND_LEARNED = 1,  /* MLP correction with diagonal fallback */
    FCI_PRECOND_OLSEN = 2     /* diagonal correction made orthogonal to the Ritz vector in the
                                 (lambda - D)^-1 metric (Olsen, Jorgensen, Simons 1990); not in the
                                 reference */
};

/* Options of FCISolver::set_options (fci_solver.cc:156-172) and of the ActiveSpaceMethod base
 * class (active_space_method.h:285-320). Defaults: fci_default_options(). */
typedef struct fci_options {
    int nroot;               /* ActiveSpaceMethod nroot_ */
    int root;                /* ROOT */
    int multiplicity;        /* StateInfo::multiplicity */
    int guess_per_root;      /* DL_GUESS_PER_ROOT */
    int ndets_per_guess;     /* DL_DETS_PER_GUESS */
    int collapse_per_root;   /* DL_COLLAPSE_PER_ROOT */
    int subspace_per_root;   /* DL_SUBSPACE_PER_ROOT */
    int maxiter;             /* DL_MAXITER */
    double e_convergence;    /* E_CONVERGENCE */
    double r_convergence;    /* R_CONVERGENCE */
    int die_if_not_converged;
    int print;               /* PrintLevel: 0 quiet .. 4 debug */
    int precision;           /* enum fci_precision */
    int preconditioner;      /* enum fci_preconditioner */
    double mixed_switch;     /* mixed precision: residual norm at which FP64 refinement starts */
    int spin_adapt;          /* CI_SPIN_ADAPT: Davidson-Liu in the CSF basis of the multiplicity */
    int spin_adapt_full_preconditioner; /* CI_SPIN_ADAPT_FULL_PRECONDITIONER: <CSF|H|CSF> diagonal */
    int ms0_symmetry;        /* no counterpart in the reference (default 1): with na == nb, an odd
                                multiplicity and equal alpha / beta integrals, keep the vectors in the
                                spin-flip-symmetric subspace C^T = (-1)^S C and build sigma without its
                                beta same-spin term (fci_sigma_ms0_device); 0 = always the full sigma */
} fci_options;

void fci_default_options(fci_options* opts);

/* FCISolver(StateInfo, nroot, MOSpaceInfo, ActiveSpaceIntegrals), fci_solver.h:63-64 and
 * fci_solver.cc:55-65. Builds the string tables and excitation lists on the device
 * (FCIStringLists::startup, fci_string_lists.cc:52-205). */
int fci_create(int n, int na, int nb, const double* oei_a, const double* oei_b,
               const double* tei_aa, const double* tei_ab, const double* tei_bb, double e_core,
               fci_solver** out);
void fci_destroy(fci_solver* s);
const char* fci_last_error(const fci_solver* s); /* also valid for s == NULL (global error) */

/* Point-group symmetry (FCISolver's StateInfo irrep and the active orbitals' irreps from
 * MOSpaceInfo, fci_solver.cc:55-65; determinant symmetry = XOR of the occupied orbitals' irreps,
 * fci_string_lists.cc:52-205): orbsym[n] in [0, 8) (Cotton order), irrep = target irrep. The solver
 * then works in the determinants of that irrep (the reference's FCIVector blocks,
 * fci_vector.cc:92-115); vectors passed to and returned by the C1 entry points (fci_evec,
 * fci_sigma*) keep the C1 layout and are zero outside the sector. */
int fci_set_symmetry(fci_solver* s, const int* orbsym, int irrep);
/* number of determinants of the target irrep (fci_ndet without symmetry) */
int64_t fci_sector_size(const fci_solver* s);
/* FCIStringLists string classes (fci_string_lists.cc:258-293, fci_string_address.cc:36-59): for
 * every string in C1 order, its irrep and its address within that irrep (all 0 / the C1 address
 * without symmetry); nstr entries each */
int fci_string_classes(fci_solver* s, int spin, int* irrep, int64_t* address);
/* eigenvector `root` in the reference's symmetry-blocked order (fci_sector_size doubles): alpha
 * irrep ha outer, then alpha address within ha, then beta address within ha ^ irrep
 * (FCIVector::copy_to, fci_vector.cc:108-169) */
int fci_evec_sector(fci_solver* s, int root, double* out);

/* GenCI's GAS restriction (genci/genci_solver.cc, genci/ci_occupation.cc:64-123, the GASn /
 * GASnMIN / GASnMAX options): ngas <= 6 spaces, gas_of_orbital[n] in [0, ngas) for every active
 * orbital, gas_min / gas_max electrons (both spins) per space. The solver then works in the
 * determinants that satisfy every bound (and the irrep of fci_set_symmetry); vectors keep the C1
 * layout and are zero outside. ngas = 0 removes the restriction. sigma here runs over all
 * determinants and is restricted afterwards (the FCI cost). */
int fci_set_gas(fci_solver* s, int ngas, const int* gas_of_orbital, const int* gas_min, const int* gas_max);
/* determinants of the space (sector and GAS restrictions; fci_ndet without either) */
int64_t fci_space_size(const fci_solver* s);

/* FCISolver::set_options, fci_solver.cc:156-172 */
int fci_set_options(fci_solver* s, const fci_options* opts);
/* ActiveSpaceMethod::set_active_space_integrals, active_space_method.h:255 (keeps the
 * Davidson-Liu basis for a warm restart, fci_solver.cc:204-237) */
int fci_set_integrals(fci_solver* s, const double* oei_a, const double* oei_b,
                      const double* tei_aa, const double* tei_ab, const double* tei_bb,
                      double e_core);

/* FCISolver::compute_energy, fci_solver.cc:177-314. *energy = energy of `root`. */
int fci_compute_energy(fci_solver* s, double* energy);
/* ActiveSpaceMethod::energies() / spin2() (active_space_method.h:261-264); nroot values each */
int fci_energies(const fci_solver* s, double* energies, double* spin2);
/* Statistics of the last solve: iterations, sigma builds, FP64 sigma builds, seconds,
 * converged flag, learned-preconditioner fallback flag. */
int fci_solve_stats(const fci_solver* s, int* iterations, int* nsigma, int* nsigma_fp64,
                    double* seconds, int* converged, int* fell_back);
/* No counterpart in the reference: the spin-flip sign eta (+1 / -1) of the last solve (C^T = eta C:
 * its determinant vectors were kept in that subspace, or, spin adapted, its sigma took the
 * symmetric form), or 0 when it used the full sigma. */
int fci_solve_ms0(const fci_solver* s, int* eta);
/* FCISolver::evecs row `root` (fci_solver.cc:79-96), ndet doubles, host buffer */
int fci_evec(const fci_solver* s, int root, double* out);
int64_t fci_ndet(const fci_solver* s);
int64_t fci_nstr(const fci_solver* s, int spin);

/* FCISolver::rdms -> FCIVector::compute_rdms (fci_solver_rdms.cc:40-104,
 * fci_vector_rdm.cc:43-393). level 1: g1a,g1b (n^2); level 2 adds g2aa,g2ab,g2bb (n^4), with
 * <a+p a+q a_s a_r> stored at [p,q,r,s]. Pass NULL for arrays not wanted. */
int fci_rdms(fci_solver* s, int root_left, int root_right, int level, double* g1a, double* g1b,
             double* g2aa, double* g2ab, double* g2bb);
/* The 3-RDMs of FCIVector::compute_rdms at max_rdm_level 3 (fci_vector_rdm.cc:70-103, 395-565):
 * g3aaa, g3aab, g3abb, g3bbb (n^6 each, NULL for blocks not wanted), element [p,q,r,s,t,u] =
 * <a+p a+q a+r a_u a_t a_s> with spins aaa / (p,q,s,t alpha, r,u beta) / (p,s alpha, q,r,t,u beta) /
 * bbb. For root_left != root_right the reference's hole-string loops place the right state on the
 * creation side, i.e. <right| ... |left>; this follows them. */
int fci_rdms3(fci_solver* s, int root_left, int root_right, double* g3aaa, double* g3aab, double* g3abb,
              double* g3bbb);
/* ActiveSpaceMethod::transition_rdms (active_space_method.h:151-154) as FCISolver implements it
 * (fci_solver_transition_rdms.cc:37-99, fci_vector_transition_rdm.cc:43-163): the 1-RDMs
 * <left root_left| a+p a_q |right root_right> of each spin between the eigenvectors of two solved
 * solvers with the same orbitals and electron counts. level must be 1: levels 2 and 3 return
 * FCI_ERR_ARG with the reference's message. (The reference's beta block reads the alpha electron
 * count, fci_vector_transition_rdm.cc:112-113, and so is zero whenever na == 0; here it is
 * computed.) */
int fci_transition_rdms(fci_solver* left, int root_left, fci_solver* right, int root_right, int level,
                        double* g1a, double* g1b);
/* ---- CSF spin adaptation (CI_SPIN_ADAPT; sparse_ci/ci_spin_adaptation.cc, fci_solver.cc:135-141,
 * 240-263). The CSFs are those of the solver's multiplicity (and irrep, with fci_set_symmetry) in
 * the reference's order: configurations sorted by (open-shell mask, closed-shell mask), genealogical
 * couplings per configuration. Determinant vectors use the C1 layout. */
/* SpinAdapter::ncsf() and the number of nonzero CSF-determinant couplings */
int fci_csf_size(fci_solver* s, int64_t* ncsf, int64_t* ncoupling);
/* SpinAdapter::det_C_to_csf_C / csf_C_to_det_C (ci_spin_adaptation.cc:89-121), host buffers */
int fci_det_to_csf(fci_solver* s, const double* det, double* csf);
int fci_csf_to_det(fci_solver* s, const double* csf, double* det);
/* FCISolver::form_Hdiag_csf (fci_solver_initial_guess.cc:111-149), honouring
 * spin_adapt_full_preconditioner; ncsf doubles */
int fci_hdiag_csf(fci_solver* s, double* out);
/* FCISolver::evecs() row `root` of a spin-adapted solve: the eigenvector in the CSF basis */
int fci_evec_csf(const fci_solver* s, int root, double* out);

/* FCIVector::compute_spin2 (fci_vector_spin2.cc:38-75) of a host vector */
int fci_spin2_of(fci_solver* s, const double* C, double* spin2);

/* FCIVector::Hamiltonian (fci_vector_hamiltonian.cc:46-82): sigma = H C.
 * Host buffers, FP64. */
int fci_sigma(fci_solver* s, const double* C, double* sigma);
/* Same on device buffers already resident in HBM (no copies). dtype 0 = FP64, 1 = FP32
 * (mixed-precision kernels). stream may be NULL (default stream). */
int fci_sigma_device(fci_solver* s, const void* C_dev, void* sigma_dev, int dtype, void* stream);
/* No counterpart in the reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). fci_sigma_device with
 * its scratch (the transposed C and its product, the same-spin intermediate) taken from one
 * static device array of FCI_NOALLOC_SCRATCH_BYTES per GPU instead of buffers the solver
 * allocates: no allocation, no host synchronization, no timings. Problems that need more scratch
 * are refused with FCI_ERR_ARG before anything is launched (CAS(12,12) and smaller fit in both
 * precisions, CAS(14,14) does not). The solver must have been created on the current device.
 * Every call on a GPU shares that GPU's array and leaves its contents behind: calls that
 * overlap on one GPU (other streams or threads) race. */
#define FCI_NOALLOC_SCRATCH_BYTES (64u << 20)
int fci_sigma_device_noalloc(fci_solver* s, const void* C_dev, void* sigma_dev, int dtype, void* stream);
/* No counterpart in the reference: sigma = H C for an Ms = 0 vector with C^T = eta C (eta = +1 or
 * -1; states of spin S have eta = (-1)^S in this determinant order), on device buffers as
 * fci_sigma_device. The beta same-spin term is eta times the transposed alpha one and is not built;
 * the alpha-beta term enters through its eta-symmetric part, so sigma is exactly eta-symmetric.
 * Needs na == nb and equal alpha and beta integrals (FCI_ERR_ARG otherwise); C must be
 * eta-symmetric (not checked: for other C the result is not H C). */
int fci_sigma_ms0_device(fci_solver* s, const void* C_dev, void* sigma_dev, int dtype, int eta, void* stream);
/* Per-kernel device time (ms) of the last fci_sigma_device call measured with HIP events:
 * [0] same-spin beta (incl. transposes), [1] same-spin alpha, [2] alpha-beta. */
int fci_sigma_timings(const fci_solver* s, double* ms3);
/* FCISolver::form_Hdiag_det (fci_solver_initial_guess.cc:151-172), host buffer of ndet */
int fci_hdiag(fci_solver* s, double* out);

/* String tables, built on the device, for parity checks against FCIStringLists. */
/* make_fci_strings (fci_string_lists.cc:258-293): nstr bit patterns in address order */
int fci_strings(fci_solver* s, int spin, uint64_t* out);
/* get_alfa_vo_list / get_beta_vo_list (fci_string_vo_list.cc:47-66): count, then entries */
int64_t fci_vo_size(fci_solver* s, int spin, int p, int q);
int fci_vo_list(fci_solver* s, int spin, int p, int q, double* sign, uint32_t* I, uint32_t* J);
/* get_*_oo_list (fci_string_oo_list.cc:40-57) for pair index pq (make_pair_list order) */
int64_t fci_oo_size(fci_solver* s, int spin, int pq);
int fci_oo_list(fci_solver* s, int spin, int pq, double* sign, uint32_t* I, uint32_t* J);
/* get_*_vvoo_list (fci_string_vvoo_list.cc:37-54) for p>q, r>s, (p,q) != (r,s) */
int64_t fci_vvoo_size(fci_solver* s, int spin, int p, int q, int r, int t);
int fci_vvoo_list(fci_solver* s, int spin, int p, int q, int r, int t, double* sign,
                  uint32_t* I, uint32_t* J);

/* ---- Davidson-Liu solver on its own (forte/helpers/davidson_liu_solver.h:52-240 and the
 * test bindings of forte/api/dl_solver_api.cc:45-86). Vectors live on the device. ---------- */
int fci_dl_create(int64_t size, int nroot, int collapse_per_root, int subspace_per_root,
                  fci_dl** out);
void fci_dl_destroy(fci_dl* d);
const char* fci_dl_last_error(const fci_dl* d);
int fci_dl_add_h_diag(fci_dl* d, const double* h_diag, int64_t size);
/* sparse vectors: n vectors, entries [off[i], off[i+1]) of (idx, val) */
int fci_dl_add_guesses(fci_dl* d, int n, const int64_t* off, const int64_t* idx, const double* val);
int fci_dl_add_project_out_vectors(fci_dl* d, int n, const int64_t* off, const int64_t* idx,
                                   const double* val);
/* add_test_sigma_builder(M): dense size x size row-major matrix, applied on the device */
int fci_dl_set_dense_sigma(fci_dl* d, const double* M);
int fci_dl_set_e_convergence(fci_dl* d, double v);
int fci_dl_set_r_convergence(fci_dl* d, double v);
int fci_dl_set_maxiter(fci_dl* d, int v);
int fci_dl_set_print_level(fci_dl* d, int v);
int fci_dl_reset(fci_dl* d);
int fci_dl_solve(fci_dl* d, int* converged);
int fci_dl_eigenvalues(const fci_dl* d, double* out); /* nroot values */
int fci_dl_eigenvector(const fci_dl* d, int k, double* out);

/* ---- device buffers (for callers without their own device allocator) --------------------- */
int fci_device_alloc(size_t bytes, void** out);
int fci_device_free(void* p);
int fci_device_upload(void* dst_dev, const void* src_host, size_t bytes);
int fci_device_download(void* dst_host, const void* src_dev, size_t bytes);
int fci_device_synchronize(void);
/* device-to-device copy of `bytes` (a plain copy kernel; used to calibrate traffic counters) */
int fci_device_copy(void* dst_dev, const void* src_dev, size_t bytes);
/* Time `nrep` sigma builds on device buffers with HIP events on one stream: ms_per_sigma is the
 * mean wall time per build; ms3 the mean per-kernel-group device time (as fci_sigma_timings). */
int fci_sigma_bench(fci_solver* s, const void* C_dev, void* sigma_dev, int dtype, int nrep,
                    double* ms_per_sigma, double* ms3);

/* ---- learned preconditioner ---------------------------------------------------------------
 * No counterpart in the reference: an MLP (34 -> 128 -> 128 -> 1, GELU) that scales Forte's
 * diagonal Davidson correction r / (lambda - H_II) (davidson_liu_solver.cc:432-448) element-wise
 * by exp(MLP(x)); y = 0 reproduces the reference. Parameters are FCI_LP_NPARAM floats in the
 * order W1[128][34], b1[128], W2[128][128], b2[128], w3[128], b3. Used when fci_options
 * .preconditioner = FCI_PRECOND_LEARNED; the diagonal preconditioner takes over for the rest of a
 * solve after two iterations without a new lowest residual norm (fci_solve_stats: fell_back). */
#define FCI_LP_NPARAM 21121
#define FCI_LP_NFEAT 34
int fci_set_lp_weights(fci_solver* s, const float* params);
/* out = learned correction of the raw residual r (host, ndet) at eigenvalue estimate lambda,
 * computed by the BF16 inference kernel with the solver's H diagonal. */
int fci_lp_precondition(fci_solver* s, const double* r, double lambda, double* out);
/* features X[m][34] of determinants idx[m] for residual r (host, ndet) and lambda */
int fci_lp_features(fci_solver* s, const int64_t* idx, int64_t m, double lambda, const double* r, float* X);
/* Training data from an FP64 solve with the diagonal preconditioner (current options): for each of
 * the first max_groups iterations, per_iter random determinants (SplitMix64 stream from seed) give
 * X (per_iter x 34), d0 = r / (lambda - H) and e = c / (c.x) - x (x: that iteration's Ritz vector
 * of root 0, c: the converged eigenvector). Buffers hold max_groups * per_iter samples; *ngroups
 * receives the number written. */
int fci_lp_harvest(fci_solver* s, int64_t per_iter, int64_t max_groups, uint64_t seed, float* X, float* d0,
                   float* e, int64_t* ngroups);
/* FP32 device MLP of the training path: y[m] = MLP(X) */
int fci_lp_forward(const float* params, const float* X, int64_t m, float* y);
/* loss = 1 - (d.e)^2 / ((d.d)(e.e)) + reg mean(y^2) with d = d0 exp(y), and d loss / d params */
int fci_lp_loss_grad(const float* params, const float* X, const float* d0, const float* e, int64_t m, float reg,
                     float* loss, float* grad);
/* Adam training on the device over ngroups groups of `group` samples (one group per step, groups
 * shuffled each epoch from seed); params is read and overwritten; epoch_loss[epochs] receives the
 * mean loss of each epoch. */
int fci_lp_train(float* params, const float* X, const float* d0, const float* e, int64_t ngroups, int64_t group,
                 int epochs, float lr, float reg, uint64_t seed, float* epoch_loss);

/* ---- many independent problems in one call ------------------------------------------------
 * The repeated-solve pattern of CASSCF micro-iterations and geometry scans (mcscf_2step.cc:187-198:
 * set integrals, FCISolver::compute_energy, fci_solver.cc:177-314), for nprob independent problems
 * at once. Problem i has norb[i] orbitals, nalpha[i] / nbeta[i] electrons and integrals in Forte's
 * layout (as fci_create). Problems of one shape share string tables and advance together through
 * Davidson-Liu iterations (davidson_liu_solver.cc:180-273) that run entirely on the device: one
 * captured graph replay per iteration, no host decision inside an iteration. opts: nroot (<= 4),
 * multiplicity, guess/collapse/subspace per root (subspace * nroot <= 32), maxiter, e/r
 * convergence; FP64, diagonal preconditioner. Outputs: energies[i * nroot + k]; optional
 * iterations[i], converged[i] (1/0), evecs[i] (host buffers of nroot * ndet_i doubles, or NULL),
 * stats. Errors: negative status, message from fci_batch_last_error(). */
#define FCI_BATCH_MAX_CLASSES 16
typedef struct fci_batch_stats {
    double setup_seconds;    /* tables, integrals, diagonals, guesses, graph capture */
    double iterate_seconds;  /* Davidson-Liu iterations (graph replays), autotuning included */
    int nclasses, converged, iterations, sigma_builds;
    int class_norb[FCI_BATCH_MAX_CLASSES], class_nprob[FCI_BATCH_MAX_CLASSES];
    int class_chunk[FCI_BATCH_MAX_CLASSES];          /* autotuned streaming chunk (elements) */
    int class_subbatch[FCI_BATCH_MAX_CLASSES];       /* autotuned sigma sub-batch (items) */
    int class_graph_launches[FCI_BATCH_MAX_CLASSES];
    double class_ms_per_iteration[FCI_BATCH_MAX_CLASSES]; /* timed during autotuning */
} fci_batch_stats;
int fci_batch_solve(int nprob, const int* norb, const int* nalpha, const int* nbeta,
                    const double* const* oei_a, const double* const* oei_b, const double* const* tei_aa,
                    const double* const* tei_ab, const double* const* tei_bb, const double* e_core,
                    const fci_options* opts, double* energies, int* iterations, int* converged,
                    double* const* evecs, fci_batch_stats* stats);
const char* fci_batch_last_error(void);

/* ---- seeded integrals of the benchmark configurations (host) ----------------------------- */
/* h (n*n) and chemist ERIs (pq|rs) (n^4) from a SplitMix64 stream; see fci_solver.hip. */
void fci_seeded_integrals(int n, uint64_t seed, double* h, double* eri);

#ifdef __cplusplus
}
#endif
#endif
