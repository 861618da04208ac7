// fci_internal.h -- shared declarations of the gfx950 FCI library (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace fci {

// ---- errors ---------------------------------------------------------------------------------
struct Error : std::runtime_error {
    int status;
    Error(int st, const std::string& m) : std::runtime_error(m), status(st) {}
};
[[noreturn]] void hip_fail(hipError_t e, const char* what, const char* file, int line);
#define HIPCHK(x)                                                                                \
    do {                                                                                         \
        hipError_t e_ = (x);                                                                     \
        if (e_ != hipSuccess) ::fci::hip_fail(e_, #x, __FILE__, __LINE__);                       \
    } while (0)

// ---- device buffer ---------------------------------------------------------------------------
template <class T> struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DBuf() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) HIPCHK(hipMalloc(&p, count * sizeof(T)));
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void upload(const T* h, size_t count) { HIPCHK(hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice)); }
    void download(T* h, size_t count) const { HIPCHK(hipMemcpy(h, p, count * sizeof(T), hipMemcpyDeviceToHost)); }
    std::vector<T> to_host() const {
        std::vector<T> v(n);
        if (n) download(v.data(), n);
        return v;
    }
    T* get() const { return p; }
};

// ---- combinatorics ---------------------------------------------------------------------------
constexpr int kMaxOrb = 22;  // strings fit in 32 bits; addresses in 21 bits
inline uint64_t binom_host(int n, int k) {
    if (k < 0 || k > n) return 0;
    uint64_t r = 1;
    for (int i = 1; i <= k; ++i) r = r * uint64_t(n - k + i) / uint64_t(i);
    return r;
}
// Binomial table C(a,b), a,b in [0, 32], row stride 33; device copy passed to kernels.
std::vector<uint32_t> binom_table();

// Packed table entries (u32): address in bits 0..20, a small index in bits 21..30, sign bit 31.
constexpr uint32_t kAddrMask = (1u << 21) - 1u;
constexpr uint32_t kInvalid = 0x7FFFFFFFu;

// ---- one spin: strings and the device tables the kernels read --------------------------------
struct SpinTables {
    int n = 0, k = 0;
    int64_t nstr = 0;
    DBuf<uint32_t> str;   // [nstr] occupation bits in Forte address order
    // singles into J: ns = k*(n-k) entries per J, J = sign * a+_p a_q I (p in J, q not in J);
    // packed (I | P<<21 | sign<<31), P = packed unordered pair index max(p,q)*(max+1)/2+min.
    int ns = 0;
    DBuf<uint32_t> into;  // [nstr * ns]
    // out of I along packed pair P (np = n(n+1)/2 entries per I): target J and sign, kInvalid
    // when the excitation does not exist. P = (r,r): J = I when r occupied.
    int np = 0;
    DBuf<uint32_t> out;   // [np * nstr], pair-major
    // the same targets as 16-bit entries (address | sign << 15, 0x7FFF = none) in the paired-column
    // MFMA tile layout [ceil(nstr/32)][npp/16][64 lanes][2][4] (npp = np rounded up to 16): the 8
    // targets of lane (g, c) for strings 32 it + 2c + p and pairs 16 tile + 4 g + v are one 16-byte
    // word; built only when nstr <= 32767
    int npp = 0;
    DBuf<uint16_t> outT;
    // pair cliques for the same-spin operator: hole strings K with k-d electrons (d = 2, or 1
    // when k == 1); members = K plus d electrons in vir(K); packed (I | m<<21 | sign<<31) where m
    // is the pair index a(a-1)/2+b (d = 2) or the orbital (d = 1).
    int d = 0;
    int64_t nK = 0;
    int M = 0;
    DBuf<uint32_t> clq;   // [nK * M]
    // owners: for each string I, the C(k,d) entries K * M + m with clq[K * M + m] addressing I
    int nown = 0;
    DBuf<uint32_t> own;   // [nstr * nown]
};

void build_spin_tables(SpinTables& t, int n, int k, const uint32_t* d_binom, hipStream_t st);
// point-group classes of the strings: irrep (XOR of the occupied orbitals' irreps, orbsym in
// [0, 8)) and address within that irrep in Forte's order; count[h] = strings of irrep h
void string_classes_device(const SpinTables& t, const std::vector<int>& orbsym, uint8_t* d_irrep, uint32_t* d_rank,
                           std::vector<int64_t>& count, hipStream_t st);

// Forte-order lists built on the device (parity checks): returns count, fills host arrays.
int64_t forte_vo_list(const SpinTables& t, const uint32_t* d_binom, int p, int q,
                      std::vector<double>* sign, std::vector<uint32_t>* I, std::vector<uint32_t>* J);
int64_t forte_oo_list(const SpinTables& t, const uint32_t* d_binom, int p, int q,
                      std::vector<double>* sign, std::vector<uint32_t>* I, std::vector<uint32_t>* J);
int64_t forte_vvoo_list(const SpinTables& t, const uint32_t* d_binom, int p, int q, int r, int s,
                        std::vector<double>* sign, std::vector<uint32_t>* I, std::vector<uint32_t>* J);

// H diagonal (form_Hdiag_det), identical summation order to ActiveSpaceIntegrals::energy.
void hdiag_device(const SpinTables& a, const SpinTables& b, const double* d_oei_a,
                  const double* d_oei_b, const double* d_tei_aa, const double* d_tei_ab,
                  const double* d_tei_bb, double e_core, double* d_out, hipStream_t st);

// k smallest (value, index) pairs of a device array, ties by index; result sorted ascending.
// ws: reduction buffers kept across calls (grow only), so repeated selections (one per problem
// of a batch) allocate nothing after the first; null for buffers of this call only
struct SelectScratch {
    DBuf<double> v[2];
    DBuf<int64_t> i[2];
};
void smallest_k(const double* d_v, int64_t n, int k, std::vector<double>& vals,
                std::vector<int64_t>& idx, hipStream_t st, SelectScratch* ws = nullptr);

// ---- integrals on the device ------------------------------------------------------------------
struct DevInts {
    int n = 0;
    double e_core = 0.0;
    DBuf<double> oei_a, oei_b, tei_aa, tei_ab, tei_bb;  // Forte layout, FP64 (Hdiag, RDM checks)
    // packed chemist ERIs V[P][Q] = (pq|rs), P,Q packed unordered pairs (np x np)
    DBuf<double> vpk;
    DBuf<float> vpk32;
    // the same scaled by 1/2 (exact): the alpha-beta term of the Ms = 0 spin-flip-symmetric sigma
    DBuf<double> vpk_h;
    DBuf<float> vpk32_h;
    // same-spin clique operator G'[m][m'] over members (pairs or orbitals), per spin
    DBuf<double> g2[2];
    DBuf<float> g2f[2];
    // mixed precision: g' without its diagonal (pass 1 on the MFMA) and, per string, the diagonal
    // sum_{pairs in occ(I)} g'[pair][pair] applied in FP32 by the gather pass (the diagonal
    // dominates |H C|, so keeping it off the BF16x3 products lowers the FP32 noise floor)
    DBuf<float> g2f_off[2];
    DBuf<float> gdiag[2];
    int gdim[2] = {0, 0};
};

// Host copy of the integrals in Forte's layout (the guess and packing work on the host).
struct HostInts {
    int n = 0;
    std::vector<double> oei_a, oei_b, tei_aa, tei_ab, tei_bb;
    double e_core = 0.0;
    size_t i4(int p, int q, int r, int s) const { return ((size_t(p) * n + q) * n + r) * n + s; }
};

// Integrals packed for the sigma kernels: vpk[P * np + Q] = (pq|rs) over packed pairs
// (np = n(n+1)/2), and per spin the same-spin clique operator g' (gd x gd, one-body term folded
// in). Throws if tei_ab lacks the permutational symmetry of real orbitals.
struct PackedInts {
    int np = 0;
    std::vector<double> vpk;
    std::vector<double> g2[2];
    int gd[2] = {0, 0};
};
PackedInts pack_integrals(const HostInts& hi, const SpinTables& sa, const SpinTables& sb);

// ---- sigma engine -----------------------------------------------------------------------------
enum class Prec { FP64 = 0, MIXED = 1 };

struct SigmaWork;  // scratch buffers

class SigmaEngine {
  public:
    SigmaEngine(const SpinTables* a, const SpinTables* b, const DevInts* ints);
    ~SigmaEngine();
    // sigma = H C on device buffers, FP64 or FP32 storage
    void apply(const double* C, double* S, hipStream_t st);
    void apply(const float* C, float* S, hipStream_t st);
    // the same with caller-provided scratch: the transposed vectors and the same-spin intermediate
    // are carved out of `ext` (ext_bytes, 256-byte aligned); nothing is allocated and nothing is
    // waited for. Throws before launching anything if ext_bytes < scratch_min_bytes(sizeof(T)).
    void apply_ext(const double* C, double* S, hipStream_t st, void* ext, size_t ext_bytes);
    void apply_ext(const float* C, float* S, hipStream_t st, void* ext, size_t ext_bytes);
    // sigma = H C for Ms = 0 vectors with C^T = eta C (eta = +-1; same alpha and beta strings and
    // integrals): the beta same-spin term is eta (alpha same-spin term)^T, so it is not built.
    // S = 1/2 e_core C + H_aa C + 1/2 H_ab C, then S <- S + eta S^T in place; the result is
    // exactly (bitwise) eta-symmetric.
    void apply_sym(const double* C, double* S, int eta, hipStream_t st);
    void apply_sym(const float* C, float* S, int eta, hipStream_t st);
    // the first half: S (holding 1/2 e_core C on entry) += H_aa C + 1/2 H_ab C; sigma = S + eta S^T
    void apply_sym_half(const double* C, double* S, int eta, hipStream_t st);
    void apply_sym_half(const float* C, float* S, int eta, hipStream_t st);
    // smallest external scratch one sigma needs with w-byte elements
    size_t scratch_min_bytes(size_t w) const;
    void timings(double* ms3) const;
    void set_timing(bool on) { timing_ = on; }

  private:
    template <class T> void apply_t(const T* C, T* S, hipStream_t st, unsigned char* ext, size_t ext_bytes);
    template <class T> void apply_sym_t(const T* C, T* S, int eta, hipStream_t st, bool init, bool finish);
    const SpinTables* a_;
    const SpinTables* b_;
    const DevInts* ints_;
    std::unique_ptr<SigmaWork> w_;
    bool timing_ = false;
    double ms_[3] = {0, 0, 0};
};

// ---- batched FP64 sigma ------------------------------------------------------------------------
// Many problems of one size (shared string tables, own integrals) in one launch per term; item i
// of a launch reads C = c[i] and writes sigma = s[i], and is idle when act[i] is null (act holds
// the same pointers as c). Each item has its own scratch: ct (C^T), yt (Y^T) and e (same-spin
// intermediate), sized by batch_sigma_scratch_elems. No e_core term (the caller adds it).
struct BatchSigmaTabs {
    int nitem = 0;
    const void* const* act = nullptr;
    const double* const* c = nullptr;
    double* const* s = nullptr;
    double* const* ct = nullptr;
    double* const* yt = nullptr;
    double* const* e = nullptr;
    const double* const* g2a = nullptr;
    const double* const* g2b = nullptr;
    const double* const* vpk = nullptr;
};
bool batch_sigma_supported(const SpinTables& a, const SpinTables& b, std::string* why);
// elements of scratch per item; *off_yt and *off_e: offsets of Y^T and E inside it (C^T at 0)
size_t batch_sigma_scratch_elems(const SpinTables& a, const SpinTables& b, size_t* off_yt, size_t* off_e);
// sets the kernel attributes batch_sigma_f64 needs (call before capturing it in a graph)
void batch_sigma_prepare(const SpinTables& a, const SpinTables& b);
void batch_sigma_f64(const SpinTables& a, const SpinTables& b, int gd_a, int gd_b, const BatchSigmaTabs& t,
                     hipStream_t st);
// the same for Ms = 0 items with C^T = eta C held in packed coordinates (ms0_pack_device): c and s
// are packed vectors, ct and yt full-size scratch (C and S), the vpk tables hold halved integrals;
// no beta same-spin term is built
void batch_sigma_f64_ms0(const SpinTables& a, const SpinTables& b, int gd_a, int eta, const BatchSigmaTabs& t,
                         hipStream_t st);

// per-string diagonal of the same-spin clique operator: out[I] = sum over the d-subsets P of
// occ(I) of g[P][P] (d = 2: pairs a > b, index a(a-1)/2 + b; d = 1: orbitals)
void clique_diagonal_device(const SpinTables& sp, const double* d_g, int gdim, float* d_out, hipStream_t st);

// transposes: out[c*R + r] = in[r*Cc + c] (+ optional accumulate of beta * base)
template <class T>
void transpose_device(const T* in, T* out, int64_t rows, int64_t cols, hipStream_t st);
template <class T>
void transpose_add_device(const T* in, T* out, int64_t rows, int64_t cols, const T* base,
                          T base_scale, hipStream_t st);
// square N x N matrix in place: S <- S + eta S^T (eta = +-1; each mirrored pair is computed once,
// so the result is exactly eta-symmetric); with avg, S <- (S + eta S^T) / 2
template <class T>
void sym_add_device(T* S, int64_t N, int eta, bool avg, hipStream_t st);
// Ms = 0 vectors with C^T = eta C in packed coordinates: the upper triangle of C[Ia][Ib] (N x N),
// entry (i, j), i < j, = (C_ij + eta C_ji) / sqrt2, entry (i, i) = C_ii (eta = +1 only; eta = -1
// vectors have a zero diagonal and no diagonal entries). These are the coordinates in an
// orthonormal basis of the eta-symmetric subspace, so dot products and norms are preserved.
int64_t ms0_packed_size(int64_t N, int eta);
int64_t ms0_packed_index(int64_t i, int64_t j, int64_t N, int eta);
// a sparse determinant vector (index Ia * N + Ib) in packed coordinates: pair (Ia, Ib), Ia < Ib ->
// (c + eta c') / sqrt2, diagonal (eta = +1) -> c; later duplicates of an index overwrite
std::vector<std::pair<int64_t, double>> ms0_pack_sparse(const std::vector<std::pair<int64_t, double>>& v,
                                                        int64_t N, int eta);
// packed(i, j) = s_off (F_ij + mirror F_ji) for i < j, s_diag F_ii on the diagonal (vectors:
// mirror = eta, s_off = 1/sqrt2, s_diag = 1; a symmetric array such as the H diagonal: mirror = +1,
// s_off = 1/2, s_diag = 1; the packed sigma of the first half S of the symmetric form, sigma =
// S + eta S^T: mirror = eta, s_off = sqrt2, s_diag = 2)
template <class TI, class TO>
void ms0_pack_device(const TI* full, TO* packed, int64_t N, int eta, int mirror, double s_off, double s_diag,
                     hipStream_t st);
// the eta-symmetric full vector of packed coordinates (exactly symmetric in the output precision);
// with s_init, also s_init = s_scale * full
template <class TI, class TO>
void ms0_unpack_device(const TI* packed, TO* full, int64_t N, int eta, hipStream_t st, TO* s_init = nullptr,
                       double s_scale = 0.0);
void ms0_pack_mask_device(const uint8_t* full, uint8_t* packed, int64_t N, int eta, hipStream_t st);

// ---- CSF spin adaptation (fci_csf.hip) -----------------------------------------------------------
// a determinant as its alpha / beta occupation bits
struct Det {
    uint32_t a, b;
    bool operator==(const Det& o) const { return a == o.a && b == o.b; }
};

// configurations with N open shells: their masks, CSF offsets and the shared table T_N
struct CsfClass {
    int N = 0, ncsf = 0, ndp = 0;  // open shells, CSFs and determinant patterns per configuration
    int64_t nconf = 0;
    // host copies: docc (bit 31: the configuration's sign), socc, CSF offset; pattern masks
    // (bit k: k-th lowest open shell beta); T_N by rows (CSF) and by columns (pattern)
    std::vector<uint32_t> h_docc, h_socc, h_occ, h_rj, h_ci;
    std::vector<int64_t> h_off;
    std::vector<int> h_rp, h_cp;
    std::vector<double> h_rv, h_cv;
    DBuf<uint32_t> docc, socc, occ, rj, ci;
    DBuf<int64_t> off;
    DBuf<int> rp, cp;
    DBuf<double> rv, cv;
    DBuf<double> td;  // T_N dense [ncsf][ndp], for classes small enough to stage it in LDS
};

// GAS restriction: orbital mask of each space and its electron bounds (both spins)
struct GasSpec {
    std::vector<uint32_t> orbs;
    std::vector<int> gmin, gmax;
};

// Forte's SpinAdapter (sparse_ci/ci_spin_adaptation.cc) over the C1 determinant layout
// C[Ia * nbs + Ib]: CSFs of spin S (twoS = 2S) in the reference's order, optionally restricted to
// the determinants of one irrep and to the configurations a GAS restriction allows
class CsfSpace {
  public:
    CsfSpace(int n, int na, int nb, int twoS, const std::vector<int>* orbsym, int irrep, const GasSpec* gas,
             const uint32_t* d_binom, int64_t nbs);
    int64_t ncsf() const { return ncsf_; }
    int64_t ncoupling() const { return ncoupling_; }
    int64_t nconf() const { return nconf_; }
    int twoS() const { return twoS_; }
    // det_C_to_csf_C (ci_spin_adaptation.cc:89-104); TI = TO = double or float
    template <class TI, class TO> void to_csf(const TI* det, TO* csf, hipStream_t st) const;
    // csf_C_to_det_C (ci_spin_adaptation.cc:106-121); determinants outside the CSFs are zeroed
    template <class TO> void to_det(const double* csf, TO* det, int64_t ndet, hipStream_t st) const;
    // sum_j c_j^2 H_jj per CSF
    void hdiag(const double* hdet, double* hcsf, hipStream_t st) const;
    // host walk over every CSF: (CSF index, its determinants with coefficients)
    void for_each_csf(const std::function<void(int64_t, const std::vector<std::pair<Det, double>>&)>& f) const;

  private:
    int n_, na_, nb_, twoS_;
    int64_t nbs_;
    const uint32_t* binom_;
    int64_t ncsf_ = 0, ncoupling_ = 0, nconf_ = 0;
    int64_t ndet_cov_ = 0;    // determinants of the configurations (to_det writes each once)
    DBuf<uint16_t> rtab_;     // byte tables of the string ranks (n <= 16), else empty
    std::vector<CsfClass> cls_;
};

// ---- RDMs ---------------------------------------------------------------------------------------
void rdm1_device(const SpinTables& a, const SpinTables& b, const double* L, const double* R,
                 double* g1a, double* g1b, hipStream_t st);
void rdm2_same_device(const SpinTables& a, const SpinTables& b, const double* L, const double* R,
                      double* g2aa, double* g2bb, hipStream_t st);
void rdm2_ab_device(const SpinTables& a, const SpinTables& b, const double* L, const double* R,
                    double* g2ab, hipStream_t st);
double spin2_device(const SpinTables& a, const SpinTables& b, const double* L, const double* R,
                    hipStream_t st);
// compact 3-RDM blocks, rows of L/R = strings of the first spin (na1 electrons), columns = strings
// of the second spin (na2): c3 [C(n,3)][C(n,3)] over sorted triples of the first spin, c21
// [C(n,2) n][C(n,2) n] over (sorted pair of the first spin, orbital of the second); NULL to skip
void rdm3_compact_device(int n, int na1, int na2, const double* L, const double* R, int64_t nrow, int64_t ncol,
                         double* c3, double* c21, hipStream_t st);

}  // namespace fci
