// fci_csf.hip -- CSF spin adaptation on the device: FCISolver's CI_SPIN_ADAPT path
// (forte/fci/fci_solver.cc:135-141, 240-263) with Forte's SpinAdapter
// (forte/sparse_ci/ci_spin_adaptation.cc).
//
// A configuration (closed shells docc, open shells socc) with N open shells owns the CSFs of the
// genealogical spin couplings of N electrons and the determinants of the occupation patterns of
// its open shells; the coefficient of determinant pattern j in coupling i is the same
// Clebsch-Gordan product for every configuration with N open shells, times the sign of creating
// the open-shell electrons (highest first) on top of the closed shells. That sign splits into a
// factor of the configuration, (-1)^(sum over open shells p of the closed shells below p), and a
// factor of the pattern alone, (-1)^(sum over beta open shells of |docc| + alpha open shells
// above), since |docc| is fixed by N. So each class of configurations with N open shells needs one
// small table T_N[i][j] (pattern sign folded in), and a transform is, per configuration, a sparse
// product with T_N on determinants whose addresses are computed on the fly from the
// configuration's masks: no coupling list the size of the CSF space (2.3e9 couplings at
// CAS(16,16) singlet) is ever stored.
#include "fci_internal.h"

#include <algorithm>
#include <cmath>

namespace fci {

namespace {

// colex rank of the bit-reversed occupation (the string addresses of the tables)
__device__ __forceinline__ uint32_t csf_rank(uint32_t s, int n, const uint32_t* __restrict__ B) {
    uint32_t r = 0;
    int t = 0;
    while (s) {
        int i = 31 - __clz(s);
        s &= ~(1u << i);
        ++t;
        r += B[(n - 1 - i) * 33 + t];
    }
    return r;
}

// alpha / beta strings of the configuration's determinant with open-shell spin pattern m
// (bit k of m set: the k-th lowest open shell holds a beta electron)
__device__ __forceinline__ void csf_strings(uint32_t docc, uint32_t socc, uint32_t m, uint32_t& a, uint32_t& b) {
    a = docc;
    b = docc;
    int k = 0;
    while (socc) {
        const uint32_t bit = socc & (0u - socc);
        socc ^= bit;
        if ((m >> k) & 1u) b |= bit;
        else a |= bit;
        ++k;
    }
}

constexpr uint32_t kConfNeg = 0x80000000u;  // sign bit of the configuration factor, kept in docc

// det -> csf: thread per (configuration, CSF row); csf[off + i] = s_c sum_j T[i][j] det[addr(c, j)]
template <class TI, class TO>
__global__ void k_csf_gather(const TI* __restrict__ det, TO* __restrict__ csf, const uint32_t* __restrict__ docc,
                             const uint32_t* __restrict__ socc, const int64_t* __restrict__ off, int64_t nconf,
                             int ncsf, const int* __restrict__ rp, const uint32_t* __restrict__ rj,
                             const double* __restrict__ rv, const uint32_t* __restrict__ occ, int n, int64_t nbs,
                             const uint32_t* __restrict__ B) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nconf * ncsf) return;
    const int64_t c = t / ncsf;
    const int i = int(t % ncsf);
    const uint32_t dw = docc[c], so = socc[c];
    const uint32_t d = dw & ~kConfNeg;
    double acc = 0.0;
    for (int e = rp[i]; e < rp[i + 1]; ++e) {
        uint32_t a, b;
        csf_strings(d, so, occ[rj[e]], a, b);
        const int64_t addr = int64_t(csf_rank(a, n, B)) * nbs + csf_rank(b, n, B);
        acc += rv[e] * double(det[addr]);
    }
    csf[off[c] + i] = TO((dw & kConfNeg) ? -acc : acc);
}

// csf -> det: thread per (configuration, determinant pattern); every determinant of the sector
// belongs to exactly one configuration, so each is written once
template <class TO>
__global__ void k_csf_scatter(const double* __restrict__ csf, TO* __restrict__ det, const uint32_t* __restrict__ docc,
                              const uint32_t* __restrict__ socc, const int64_t* __restrict__ off, int64_t nconf,
                              int ndp, const int* __restrict__ cp, const uint32_t* __restrict__ ci,
                              const double* __restrict__ cv, const uint32_t* __restrict__ occ, int n, int64_t nbs,
                              const uint32_t* __restrict__ B) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nconf * ndp) return;
    const int64_t c = t / ndp;
    const int j = int(t % ndp);
    const uint32_t dw = docc[c], so = socc[c];
    const int64_t o = off[c];
    double acc = 0.0;
    for (int e = cp[j]; e < cp[j + 1]; ++e) acc += cv[e] * csf[o + ci[e]];
    uint32_t a, b;
    csf_strings(dw & ~kConfNeg, so, occ[j], a, b);
    det[int64_t(csf_rank(a, n, B)) * nbs + csf_rank(b, n, B)] = TO((dw & kConfNeg) ? -acc : acc);
}

// FCISolver::form_Hdiag_csf without the full preconditioner (fci_solver_initial_guess.cc:134-141):
// sum_j T[i][j]^2 H_jj
__global__ void k_csf_hdiag(const double* __restrict__ hd, double* __restrict__ out, const uint32_t* __restrict__ docc,
                            const uint32_t* __restrict__ socc, const int64_t* __restrict__ off, int64_t nconf,
                            int ncsf, const int* __restrict__ rp, const uint32_t* __restrict__ rj,
                            const double* __restrict__ rv, const uint32_t* __restrict__ occ, int n, int64_t nbs,
                            const uint32_t* __restrict__ B) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nconf * ncsf) return;
    const int64_t c = t / ncsf;
    const int i = int(t % ncsf);
    const uint32_t d = docc[c] & ~kConfNeg, so = socc[c];
    double acc = 0.0;
    for (int e = rp[i]; e < rp[i + 1]; ++e) {
        uint32_t a, b;
        csf_strings(d, so, occ[rj[e]], a, b);
        acc += rv[e] * rv[e] * hd[int64_t(csf_rank(a, n, B)) * nbs + csf_rank(b, n, B)];
    }
    out[off[c] + i] = acc;
}

inline unsigned grid_of(int64_t work) { return unsigned((work + 255) / 256); }

// all 0/1 sequences of length N with `ones` ones, lexicographic (0 before 1), optionally with
// position 0 held at 0: what std::next_permutation enumerates from the sorted start
std::vector<std::vector<int>> ordered_01(int N, int ones, bool first_zero) {
    std::vector<std::vector<int>> out;
    std::vector<int> s(N, 0);
    for (int k = N - ones; k < N; ++k) s[k] = 1;
    if (first_zero && N > 0 && s[0] == 1) return out;  // no sequence with a leading 0 and that many ones
    const int b = first_zero ? 1 : 0;
    do {
        out.push_back(s);
    } while (N > b && std::next_permutation(s.begin() + b, s.end()));
    return out;
}

// Clebsch-Gordan factor of one coupling step (ci_spin_adaptation.cc:46-58): twoS, twoM after the
// step, dS = +-1 the spin step, dM = +-1 the projection step
double cg_step(int twoS, int twoM, int dS, int dM) {
    if (dS == 1) return std::sqrt(0.5 * (twoS + dM * twoM) / twoS);
    return -dM * std::sqrt(0.5 * (twoS + 2.0 - dM * twoM) / (twoS + 2.0));
}

}  // namespace

CsfSpace::CsfSpace(int n, int na, int nb, int twoS, const std::vector<int>* orbsym, int irrep, const uint32_t* d_binom,
                   int64_t nbs)
    : n_(n), na_(na), nb_(nb), twoS_(twoS), nbs_(nbs), binom_(d_binom) {
    const int twoMs = na - nb;
    if (twoS < 0 || twoS < std::abs(twoMs) || (twoS - twoMs) % 2)
        throw Error(-1, "spin adaptation: the multiplicity is not compatible with the electron counts");
    // configurations: closed shells (d of them) and N = na + nb - 2d open shells, sorted by
    // (socc, docc) as integers (the order of std::set<Configuration>, bitarray.hpp:301-326), those
    // with fewer than 2S open shells dropped (ci_spin_adaptation.cc:151-156)
    std::vector<uint64_t> keys;
    auto subsets = [](int nn, int k) {
        std::vector<uint32_t> v;
        if (k > nn || k < 0) return v;
        if (k == 0) {
            v.push_back(0u);
            return v;
        }
        uint32_t s = (1u << k) - 1u;
        const uint32_t lim = 1u << nn;
        while (s < lim) {
            v.push_back(s);
            const uint32_t c = s & (0u - s), r = s + c;
            s = (((r ^ s) >> 2) / c) | r;
        }
        return v;
    };
    for (int d = 0; d <= std::min(na, nb); ++d) {
        const int N = na + nb - 2 * d;
        if (N < twoS || N < std::abs(twoMs) || d + N > n) continue;
        for (uint32_t so : subsets(n, N)) {
            if (orbsym) {
                int h = 0;
                for (int p = 0; p < n; ++p)
                    if ((so >> p) & 1u) h ^= (*orbsym)[p];
                if (h != irrep) continue;
            }
            // closed shells: d-subsets of the remaining orbitals
            std::vector<int> rest;
            for (int p = 0; p < n; ++p)
                if (!((so >> p) & 1u)) rest.push_back(p);
            for (uint32_t sub : subsets(int(rest.size()), d)) {
                uint32_t dm = 0;
                for (size_t q = 0; q < rest.size(); ++q)
                    if ((sub >> q) & 1u) dm |= 1u << rest[q];
                keys.push_back((uint64_t(so) << 32) | dm);
            }
        }
    }
    std::sort(keys.begin(), keys.end());
    // classes by open-shell count: couplings, patterns and the table T_N
    std::vector<int> cls_of(n + 1, -1);
    for (uint64_t k : keys) {
        const int N = __builtin_popcount(uint32_t(k >> 32));
        if (cls_of[N] >= 0) continue;
        cls_of[N] = int(cls_.size());
        cls_.emplace_back();
        CsfClass& C = cls_.back();
        C.N = N;
        const int d = (na + nb - N) / 2;
        const int nbo = (N - twoMs) / 2;  // beta open shells
        // make_spin_couplings (ci_spin_adaptation.cc:232-258): first step up, never negative
        std::vector<std::vector<int>> cpl;
        if (N == 0) cpl.push_back({});
        else
            for (auto& s : ordered_01(N, N - (N + twoS) / 2, true)) {
                int p = 0;
                bool ok = true;
                for (int x : s) {
                    p += 1 - 2 * x;
                    if (p < 0) ok = false;
                }
                if (ok) cpl.push_back(s);
            }
        // make_determinant_occupations (:260-279)
        std::vector<std::vector<int>> pat = N == 0 ? std::vector<std::vector<int>>{{}} : ordered_01(N, nbo, false);
        C.ncsf = int(cpl.size());
        C.ndp = int(pat.size());
        C.h_occ.resize(C.ndp);
        std::vector<double> psign(C.ndp);
        for (int j = 0; j < C.ndp; ++j) {
            uint32_t m = 0;
            int e = 0, na_above = 0;
            for (int k = N - 1; k >= 0; --k) {
                if (pat[j][k]) {
                    m |= 1u << k;
                    e += d + na_above;
                } else {
                    ++na_above;
                }
            }
            C.h_occ[j] = m;
            psign[j] = (e & 1) ? -1.0 : 1.0;
        }
        // overlaps (ci_spin_adaptation.cc:60-81), nonzero ones only, in pattern order per coupling
        C.h_rp.assign(1, 0);
        for (const auto& cp : cpl) {
            for (int j = 0; j < C.ndp; ++j) {
                double ov = 1.0;
                int p = 0, q = 0;
                for (int i = 0; i < N && ov != 0.0; ++i) {
                    const int dS = 1 - 2 * cp[i], dM = 1 - 2 * pat[j][i];
                    p += dS;
                    q += dM;
                    if (std::abs(q) > p) ov = 0.0;
                    else ov *= cg_step(p, q, dS, dM);
                }
                if (ov != 0.0) {
                    C.h_rj.push_back(uint32_t(j));
                    C.h_rv.push_back(ov * psign[j]);
                }
            }
            C.h_rp.push_back(int(C.h_rj.size()));
        }
        // columns (pattern j): its nonzeros over the couplings
        C.h_cp.assign(C.ndp + 1, 0);
        for (uint32_t j : C.h_rj) ++C.h_cp[j + 1];
        for (int j = 0; j < C.ndp; ++j) C.h_cp[j + 1] += C.h_cp[j];
        C.h_ci.resize(C.h_rj.size());
        C.h_cv.resize(C.h_rj.size());
        std::vector<int> fill(C.h_cp.begin(), C.h_cp.end() - 1);
        for (int i = 0; i < C.ncsf; ++i)
            for (int e = C.h_rp[i]; e < C.h_rp[i + 1]; ++e) {
                const int at = fill[C.h_rj[e]]++;
                C.h_ci[at] = uint32_t(i);
                C.h_cv[at] = C.h_rv[e];
            }
    }
    // configurations into their classes, with their CSF offsets in the global order
    int64_t offc = 0;
    for (uint64_t k : keys) {
        const uint32_t so = uint32_t(k >> 32), dm = uint32_t(k);
        CsfClass& C = cls_[cls_of[__builtin_popcount(so)]];
        int e = 0;  // configuration factor: closed shells below each open shell
        for (int p = 0; p < n; ++p)
            if ((so >> p) & 1u) e += __builtin_popcount(dm & ((1u << p) - 1u));
        C.h_docc.push_back(dm | ((e & 1) ? kConfNeg : 0u));
        C.h_socc.push_back(so);
        C.h_off.push_back(offc);
        offc += C.ncsf;
        ncoupling_ += int64_t(C.h_rj.size());
    }
    ncsf_ = offc;
    nconf_ = int64_t(keys.size());
    for (auto& C : cls_) {
        C.nconf = int64_t(C.h_docc.size());
        C.docc.alloc(C.nconf);
        C.docc.upload(C.h_docc.data(), C.nconf);
        C.socc.alloc(C.nconf);
        C.socc.upload(C.h_socc.data(), C.nconf);
        C.off.alloc(C.nconf);
        C.off.upload(C.h_off.data(), C.nconf);
        C.occ.alloc(C.ndp);
        C.occ.upload(C.h_occ.data(), C.ndp);
        C.rp.alloc(C.h_rp.size());
        C.rp.upload(C.h_rp.data(), C.h_rp.size());
        C.cp.alloc(C.h_cp.size());
        C.cp.upload(C.h_cp.data(), C.h_cp.size());
        const size_t nz = std::max<size_t>(1, C.h_rj.size());
        C.rj.alloc(nz);
        C.rv.alloc(nz);
        C.ci.alloc(nz);
        C.cv.alloc(nz);
        if (!C.h_rj.empty()) {
            C.rj.upload(C.h_rj.data(), C.h_rj.size());
            C.rv.upload(C.h_rv.data(), C.h_rv.size());
            C.ci.upload(C.h_ci.data(), C.h_ci.size());
            C.cv.upload(C.h_cv.data(), C.h_cv.size());
        }
    }
}

template <class TI, class TO> void CsfSpace::to_csf(const TI* det, TO* csf, hipStream_t st) const {
    for (const auto& C : cls_) {
        if (C.nconf == 0 || C.ncsf == 0) continue;
        k_csf_gather<TI, TO><<<grid_of(C.nconf * C.ncsf), 256, 0, st>>>(
            det, csf, C.docc.p, C.socc.p, C.off.p, C.nconf, C.ncsf, C.rp.p, C.rj.p, C.rv.p, C.occ.p, n_, nbs_, binom_);
    }
    HIPCHK(hipGetLastError());
}

template <class TO> void CsfSpace::to_det(const double* csf, TO* det, int64_t ndet, hipStream_t st) const {
    HIPCHK(hipMemsetAsync(det, 0, size_t(ndet) * sizeof(TO), st));
    for (const auto& C : cls_) {
        if (C.nconf == 0) continue;
        k_csf_scatter<TO><<<grid_of(C.nconf * C.ndp), 256, 0, st>>>(
            csf, det, C.docc.p, C.socc.p, C.off.p, C.nconf, C.ndp, C.cp.p, C.ci.p, C.cv.p, C.occ.p, n_, nbs_, binom_);
    }
    HIPCHK(hipGetLastError());
}

void CsfSpace::hdiag(const double* hdet, double* hcsf, hipStream_t st) const {
    for (const auto& C : cls_) {
        if (C.nconf == 0 || C.ncsf == 0) continue;
        k_csf_hdiag<<<grid_of(C.nconf * C.ncsf), 256, 0, st>>>(hdet, hcsf, C.docc.p, C.socc.p, C.off.p, C.nconf,
                                                              C.ncsf, C.rp.p, C.rj.p, C.rv.p, C.occ.p, n_, nbs_,
                                                              binom_);
    }
    HIPCHK(hipGetLastError());
}

void CsfSpace::for_each_csf(const std::function<void(int64_t, const std::vector<std::pair<Det, double>>&)>& f) const {
    std::vector<std::pair<Det, double>> ents;
    for (const auto& C : cls_)
        for (int64_t c = 0; c < C.nconf; ++c) {
            const uint32_t dw = C.h_docc[c], d = dw & ~kConfNeg, so = C.h_socc[c];
            const double s = (dw & kConfNeg) ? -1.0 : 1.0;
            for (int i = 0; i < C.ncsf; ++i) {
                ents.clear();
                for (int e = C.h_rp[i]; e < C.h_rp[i + 1]; ++e) {
                    uint32_t a = d, b = d, m = C.h_occ[C.h_rj[e]], rest = so;
                    for (int k = 0; rest; ++k) {
                        const uint32_t bit = rest & (0u - rest);
                        rest ^= bit;
                        if ((m >> k) & 1u) b |= bit;
                        else a |= bit;
                    }
                    ents.push_back({Det{a, b}, s * C.h_rv[e]});
                }
                f(C.h_off[c] + i, ents);
            }
        }
}

template void CsfSpace::to_csf<double, double>(const double*, double*, hipStream_t) const;
template void CsfSpace::to_csf<float, float>(const float*, float*, hipStream_t) const;
template void CsfSpace::to_det<double>(const double*, double*, int64_t, hipStream_t) const;
template void CsfSpace::to_det<float>(const double*, float*, int64_t, hipStream_t) const;

}  // namespace fci
