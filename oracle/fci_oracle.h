/*
 * fci_oracle.h -- CPU restatement of Forte's FCISolver hot path (test infrastructure only).
 *
 * This library is the checker for the gfx950 FCI solver. It is linked only by the tests, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py. It restates, in plain C and
 * in FP64, the algorithm of Forte 0.4.3 for one irrep (C1):
 *
 *   strings / addresses     forte/fci/fci_string_lists.cc:258-293, fci_string_address.cc:36-59
 *   VO / OO / VVOO lists    forte/fci/fci_string_vo_list.cc:68-140, fci_string_oo_list.cc:59-115,
 *                           forte/fci/fci_string_vvoo_list.cc:56-187, fci_string_lists.cc:211-229
 *   sigma = H C             forte/fci/fci_vector_hamiltonian.cc:46-272, fci_vector.cc:239-278
 *   H diagonal              forte/fci/fci_solver_initial_guess.cc:151-172,
 *                           forte/integrals/active_space_integrals.cc:206-245
 *   Slater rules            forte/integrals/active_space_integrals.cc:247-411
 *   initial guess           forte/fci/fci_solver_initial_guess.cc:44-101,
 *                           forte/sparse_ci/sparse_initial_guess.cc:51-264,
 *                           forte/sparse_ci/determinant_functions.cc:133-194,
 *                           forte/helpers/determinant_helpers.cc:48-87, helpers/helpers.cc:233-262,
 *                           forte/sparse_ci/determinant.hpp:910-955 (spin2)
 *   Davidson-Liu            forte/helpers/davidson_liu_solver.cc:45-731
 *   FCISolver driver        forte/fci/fci_solver.cc:177-314
 *   1-/2-RDMs, <S^2>        forte/fci/fci_vector_rdm.cc:43-393, fci_vector_spin2.cc:38-75
 *
 * Independent checks (not restatements) live beside it: a brute-force second-quantized
 * Hamiltonian and brute-force RDMs built from explicit creation/annihilation operators.
 *
 * Deviation from the reference, by design: the reference draws Davidson restart vectors from
 * std::random_device (davidson_liu_solver.cc:521-536); here they come from a SplitMix64 stream
 * with a fixed seed so that runs are reproducible.
 */
#ifndef FCI_ORACLE_H
#define FCI_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

/* ---- strings ---------------------------------------------------------------------------- */
/* Number of k-of-n strings. */
uint64_t orc_binom(int n, int k);
/* Enumerate all k-of-n strings in Forte's std::next_permutation order (orbital i <-> bit i). */
int64_t orc_make_strings(int n, int k, uint64_t* out);

/* ---- context: strings, addresses and excitation lists for one problem ------------------ */
/* Integrals in Forte's layout: oei n*n row-major, tei n^4 row-major physicist
 * (tei_aa/bb antisymmetrized <pq||rs>, tei_ab = <pq|rs>). e_core = scalar + frozen + nuclear. */
orc_ctx* orc_create(int n, int na, int nb, const double* oei_a, const double* oei_b,
                    const double* tei_aa, const double* tei_ab, const double* tei_bb,
                    double e_core);
/* Strings and addresses only, no excitation lists: for orc_sigma_rows, orc_hdiag and
 * orc_dense_hamiltonian at sizes whose lists the oracle cannot afford. */
orc_ctx* orc_create_strings(int n, int na, int nb, const double* oei_a, const double* oei_b,
                            const double* tei_aa, const double* tei_ab, const double* tei_bb,
                            double e_core);
int orc_has_lists(const orc_ctx* ctx);
void orc_destroy(orc_ctx* ctx);
const char* orc_last_error(void);

int64_t orc_nstr(const orc_ctx* ctx, int spin); /* spin 0 = alpha, 1 = beta */
int64_t orc_ndet(const orc_ctx* ctx);
void orc_strings(const orc_ctx* ctx, int spin, uint64_t* out);
int64_t orc_address(const orc_ctx* ctx, int spin, uint64_t str); /* -1 if not a valid string */

/* VO list (p,q): entries (sign, I, J) with J = sign a+_p a_q I, in Forte order. */
int64_t orc_vo_size(const orc_ctx* ctx, int spin, int p, int q);
void orc_vo_get(const orc_ctx* ctx, int spin, int p, int q, double* sign, uint32_t* I,
                uint32_t* J);
/* OO list for pair index pq (pairs p>q ordered as in make_pair_list). */
int64_t orc_npair(const orc_ctx* ctx);
void orc_pair(const orc_ctx* ctx, int64_t pq, int* p, int* q);
int64_t orc_oo_size(const orc_ctx* ctx, int spin, int64_t pq);
void orc_oo_get(const orc_ctx* ctx, int spin, int64_t pq, double* sign, uint32_t* I,
                uint32_t* J);
/* VVOO list (p,q,r,s), p>q, r>s, (p,q) != (r,s): J = sign a+_p a+_q a_s a_r I. */
int64_t orc_vvoo_size(const orc_ctx* ctx, int spin, int p, int q, int r, int s);
void orc_vvoo_get(const orc_ctx* ctx, int spin, int p, int q, int r, int s, double* sign,
                  uint32_t* I, uint32_t* J);

/* ---- sigma, diagonal, spin, RDMs -------------------------------------------------------- */
enum {
    ORC_H0 = 1,
    ORC_H1A = 2,
    ORC_H1B = 4,
    ORC_H2AABB = 8,
    ORC_H2AAAA = 16,
    ORC_H2BBBB = 32,
    ORC_ALL = 63
};
/* sigma = H C restricted to the terms in mask, with Forte's loop structure. */
void orc_sigma(orc_ctx* ctx, const double* C, double* sigma, int mask);
/* Per-term wall times (seconds) of the last orc_sigma call: H0,H1a,H1b,H2aabb,H2aaaa,H2bbbb */
void orc_sigma_timers(const orc_ctx* ctx, double* t6);
/* H2_aabb restricted to the beta pairs (r,s) with rs_begin <= r*n+s < rs_end (sampling). */
void orc_sigma_aabb_range(orc_ctx* ctx, const double* C, double* sigma, int rs_begin, int rs_end);
/* Timing sample of the Forte-structured sigma: every term runs only the outer-loop indices i
 * with i % stride == 0 (H1: p*n+q, H2_aaaa2: pair pq, H2_aabb: r*n+s). t6 gets the sampled
 * seconds per term (H0,H1a,H1b,H2aabb,H2aaaa,H2bbbb) and w6 the fraction of that term's list
 * entries the sample covered, so t6[i]/w6[i] extrapolates the full term. */
void orc_sigma_sample(orc_ctx* ctx, const double* C, double* sigma, int stride, double* t6, double* w6);
void orc_hdiag(const orc_ctx* ctx, double* out);
double orc_spin2(const orc_ctx* ctx, const double* C);
/* level 1: g1a,g1b (n^2); level 2 adds g2aa,g2ab,g2bb (n^4), <a+p a+q a_s a_r> at [p,q,r,s]. */
void orc_rdms(const orc_ctx* ctx, const double* CL, const double* CR, int level, double* g1a,
              double* g1b, double* g2aa, double* g2ab, double* g2bb);

/* ---- brute-force (independent) checks ---------------------------------------------------- */
/* Dense H in determinant order via Slater rules (active_space_integrals.cc:247-411). */
void orc_dense_hamiltonian(const orc_ctx* ctx, double* H);
/* Rows of sigma = H C by the Slater-Condon rules over the determinants connected to each row
 * (singles and doubles of both spins); needs only strings and addresses. */
void orc_sigma_rows(const orc_ctx* ctx, const double* C, int64_t nrows, const int64_t* rows,
                    double* out);
/* Brute-force sigma from explicit second quantization of
 * H = e_core + sum h_pq a+p a_q + 1/4 sum <pq||rs> (same spin) + sum <pq|rs> (alpha-beta). */
void orc_brute_sigma(const orc_ctx* ctx, const double* C, double* sigma);
/* Brute-force RDMs from explicit operator strings on determinants. */
void orc_brute_rdms(const orc_ctx* ctx, const double* CL, const double* CR, double* g1a,
                    double* g1b, double* g2aa, double* g2ab, double* g2bb);

/* Brute-force 3-RDMs <CL| a+p a+q a+r a_u a_t a_s |CR> at [p,q,r,s,t,u] (n^6 each) for the spin
 * patterns aaa, aab (r, u beta), abb (q, r, t, u beta) and bbb, from explicit operator strings
 * (the definition of FCIVector::test_rdms, fci_vector_rdm.cc:808-845). */
void orc_brute_rdm3(const orc_ctx* ctx, const double* CL, const double* CR, double* g3aaa, double* g3aab,
                    double* g3abb, double* g3bbb);

/* ---- dense symmetric eigensolver (cyclic Jacobi), ascending eigenvalues ---------------- */
void orc_jacobi(int n, const double* A, double* evals, double* evecs /* column-major: v_k = evecs[k*n..] */);

/* ---- Davidson-Liu on a dense matrix (the reference's add_test_sigma_builder path) ------- */
typedef struct {
    int collapse_per_root;
    int subspace_per_root;
    int maxiter;
    double e_convergence;
    double r_convergence;
} orc_dl_opts;
/* guesses: ng sparse vectors, entries [goff[g], goff[g+1]) of (gidx, gval). Same for project
 * out vectors. Returns 1 if converged, 0 if not, -1 on error (orc_last_error). evals[nroot],
 * evecs[nroot*size]. */
int orc_dl_dense(int64_t size, int nroot, const orc_dl_opts* opts, const double* M,
                 const double* hdiag, int ng, const int64_t* goff, const int64_t* gidx,
                 const double* gval, int np, const int64_t* poff, const int64_t* pidx,
                 const double* pval, double* evals, double* evecs, int* niter);

/* ---- FCISolver::compute_energy ------------------------------------------------------------ */
typedef struct {
    int nroot;
    int multiplicity;
    int guess_per_root;
    int ndets_per_guess;
    int collapse_per_root;
    int subspace_per_root;
    int maxiter;
    double e_convergence;
    double r_convergence;
} orc_fci_opts;
/* energies[nroot], spin2[nroot], evecs[nroot*ndet] (may be NULL). Returns 1/0/-1 as above. */
int orc_fci_solve(orc_ctx* ctx, const orc_fci_opts* opts, double* energies, double* spin2,
                  double* evecs, int* niter);

/* ---- seeded integral generator ----------------------------------------------------------- */
/* h (n*n), eri chemist (pq|rs) n^4; see the generator description in the C source. */
void orc_seeded_integrals(int n, uint64_t seed, double* h, double* eri);
/* Forte layout from chemist ERIs: tei_ab[p,q,r,s] = (pr|qs), tei_aa = (pr|qs) - (ps|qr). */
void orc_chem_to_forte(int n, const double* eri, double* tei_aa, double* tei_ab);

#ifdef __cplusplus
}
#endif
#endif
