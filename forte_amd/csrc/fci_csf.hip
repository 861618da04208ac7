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

// For n <= 16 the rank splits over the two bytes of the string: the high byte's bits have ordinals
// counted from the top, the low byte's start after popc(high byte), so
//   rank(s) = H[s >> 8] + L[popc(s >> 8)][s & 255]
// with H and L (9 x 256) summing the binomials of the bits of one byte value. Both are 16-bit (every
// rank at n <= 16 is below C(16, 8)) and staged in LDS by every kernel: two LDS reads per string
// instead of a chain of one dependent table load per electron.
constexpr int kRankTab = 256 * 10;
__device__ __forceinline__ void csf_rank_stage(uint16_t* rl, const uint16_t* __restrict__ g) {
    if (g)
        for (int t = threadIdx.x; t < kRankTab; t += blockDim.x) rl[t] = g[t];
}
// address of the determinant (a, b) in C[Ia * nbs + Ib]; rl: the staged byte tables (fast) or
// unused (the per-electron rank with the binomial table B)
struct CsfAddr {
    const uint16_t* rl;
    bool fast;
    int n;
    int64_t nbs;
    const uint32_t* __restrict__ B;
    __device__ __forceinline__ uint32_t rank(uint32_t s) const {
        if (fast) {
            const uint32_t h = s >> 8;
            return uint32_t(rl[h]) + uint32_t(rl[256 + (__popc(h) << 8) + (s & 255u)]);
        }
        return csf_rank(s, n, B);
    }
    __device__ __forceinline__ int64_t operator()(uint32_t a, uint32_t b) const {
        return int64_t(rank(a)) * nbs + rank(b);
    }
};

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

// Stage the determinants of configurations [0, nc) (docc, socc point at the block's first one) into
// xs[cc * ndp + j], signed by the configuration factor; rows cc in [nc, npad) are zeroed. Each
// thread computes kStageU addresses before it loads, so that many random reads are in flight per
// thread (blocks that hold large LDS tiles run few waves per CU).
constexpr int kStageU = 8;
template <class TI>
__device__ __forceinline__ void csf_stage_dets(const TI* __restrict__ det, double* xs, int nc, int npad, int ndp,
                                               const uint32_t* __restrict__ docc, const uint32_t* __restrict__ socc,
                                               const uint32_t* __restrict__ occ, const CsfAddr& addr_of) {
    const int valid = nc * ndp, total = npad * ndp;
    for (int t0 = threadIdx.x; t0 < total; t0 += kStageU * blockDim.x) {
        int64_t addr[kStageU];
        bool neg[kStageU];
#pragma unroll
        for (int u = 0; u < kStageU; ++u) {
            const int t = t0 + u * int(blockDim.x);
            addr[u] = -1;
            neg[u] = false;
            if (t < valid) {
                const int cc = t / ndp, j = t - cc * ndp;
                const uint32_t dw = docc[cc];
                uint32_t a, b;
                csf_strings(dw & ~kConfNeg, socc[cc], occ[j], a, b);
                addr[u] = addr_of(a, b);
                neg[u] = (dw & kConfNeg) != 0;
            }
        }
        double v[kStageU];
#pragma unroll
        for (int u = 0; u < kStageU; ++u) v[u] = addr[u] >= 0 ? double(det[addr[u]]) : 0.0;
#pragma unroll
        for (int u = 0; u < kStageU; ++u) {
            const int t = t0 + u * int(blockDim.x);
            if (t < total) xs[t] = neg[u] ? -v[u] : v[u];
        }
    }
}

// det -> csf: thread per (configuration, CSF row); csf[off + i] = s_c sum_j T[i][j] det[addr(c, j)]
template <class TI, class TO>
__global__ void k_csf_gather(const TI* __restrict__ det, TO* __restrict__ csf, const uint32_t* __restrict__ docc,
                             const uint32_t* __restrict__ socc, const int64_t* __restrict__ off, int64_t nconf,
                             int ncsf, const int* __restrict__ rp, const uint32_t* __restrict__ rj,
                             const double* __restrict__ rv, const uint32_t* __restrict__ occ, int n, int64_t nbs,
                             const uint32_t* __restrict__ B, const uint16_t* __restrict__ RT) {
    __shared__ uint16_t rl[kRankTab];
    const CsfAddr addr{rl, RT != nullptr, n, nbs, B};
    csf_rank_stage(rl, RT);
    __syncthreads();
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
        acc += rv[e] * double(det[addr(a, b)]);
    }
    csf[off[c] + i] = TO((dw & kConfNeg) ? -acc : acc);
}

// det -> csf with the determinants of `cpb` consecutive configurations staged in LDS (signed by
// their configuration factor), so each determinant address is computed once per block. Block
// (x, y) covers configurations [x cpb, x cpb + cpb) and CSF rows [y rpb, y rpb + rpb). The row
// table T_N is read once per row for kCsfCT configurations at a time; rows are spread by `mode`:
//   0: (group of kCsfCT configurations, row) pairs over the threads (few CSFs per configuration)
//   1: one row per thread
//   2: one row per wave, its nonzeros over the lanes (thousands per row at 12+ open shells)
constexpr int kCsfCT = 4;
__device__ __forceinline__ double csf_wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
template <class TI, class TO>
__global__ __launch_bounds__(256) void k_csf_gather_lds(const TI* __restrict__ det, TO* __restrict__ csf,
                                                        const uint32_t* __restrict__ docc,
                                                        const uint32_t* __restrict__ socc,
                                                        const int64_t* __restrict__ off, int64_t nconf, int cpb,
                                                        int rpb, int mode, int ncsf, int ndp,
                                                        const int* __restrict__ rp, const uint32_t* __restrict__ rj,
                                                        const double* __restrict__ rv, const uint32_t* __restrict__ occ,
                                                        int n, int64_t nbs, const uint32_t* __restrict__ B,
                                                        const uint16_t* __restrict__ RT) {
    __shared__ uint16_t rl[kRankTab];
    const CsfAddr addr{rl, RT != nullptr, n, nbs, B};
    csf_rank_stage(rl, RT);
    __syncthreads();
    extern __shared__ double xs[];  // [cpb][ndp]
    const int64_t c0 = int64_t(blockIdx.x) * cpb;
    const int nc = int(min(int64_t(cpb), nconf - c0));
    const int r0 = blockIdx.y * rpb, r1 = min(ncsf, r0 + rpb);
    csf_stage_dets(det, xs, nc, nc, ndp, docc + c0, socc + c0, occ, addr);
    __syncthreads();
    const int ng = (nc + kCsfCT - 1) / kCsfCT;
    if (mode == 2) {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
        for (int i = r0 + wave; i < r1; i += nw) {
            const int e0 = rp[i], e1 = rp[i + 1];
            for (int g = 0; g < ng; ++g) {
                const int cb = g * kCsfCT, kk = min(kCsfCT, nc - cb);
                const double* x0 = xs + cb * ndp;
                double acc[kCsfCT];
#pragma unroll
                for (int k = 0; k < kCsfCT; ++k) acc[k] = 0.0;
                // four chunks of 64 nonzeros per pass: their loads are in flight together
                for (int e = e0 + lane; e < e1; e += 4 * 64) {
                    double v[4];
                    uint32_t jj[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int eu = e + 64 * u;
                        v[u] = eu < e1 ? rv[eu] : 0.0;
                        jj[u] = eu < e1 ? rj[eu] : 0u;
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u)
#pragma unroll
                        for (int k = 0; k < kCsfCT; ++k)
                            if (k < kk) acc[k] += v[u] * x0[k * ndp + jj[u]];
                }
#pragma unroll
                for (int k = 0; k < kCsfCT; ++k) {
                    const double t = csf_wave_sum(acc[k]);
                    if (lane == 0 && k < kk) csf[off[c0 + cb + k] + i] = TO(t);
                }
            }
        }
        return;
    }
    // modes 0 and 1: work item = (group of configurations, row)
    const int nr = r1 - r0;
    const int items = mode == 0 ? ng * nr : nr;
    for (int t = threadIdx.x; t < items; t += blockDim.x) {
        const int i = r0 + (mode == 0 ? t % nr : t);
        const int g0 = mode == 0 ? t / nr : 0, g1 = mode == 0 ? g0 + 1 : ng;
        const int e0 = rp[i], e1 = rp[i + 1];
        for (int g = g0; g < g1; ++g) {
            const int cb = g * kCsfCT, kk = min(kCsfCT, nc - cb);
            const double* x0 = xs + cb * ndp;
            double acc[kCsfCT];
#pragma unroll
            for (int k = 0; k < kCsfCT; ++k) acc[k] = 0.0;
            for (int e = e0; e < e1; ++e) {
                const double v = rv[e];
                const uint32_t jj = rj[e];
#pragma unroll
                for (int k = 0; k < kCsfCT; ++k)
                    if (k < kk) acc[k] += v * x0[k * ndp + jj];
            }
#pragma unroll
            for (int k = 0; k < kCsfCT; ++k)
                if (k < kk) csf[off[c0 + cb + k] + i] = TO(acc[k]);
        }
    }
}

// LDS row stride of a staged dense T_N: odd, so rows read by consecutive lanes fall in distinct banks
__host__ __device__ __forceinline__ int csf_tstride(int ndp) { return ndp | 1; }
// dense T_N [ncsf][ndp] from global memory into LDS rows of stride ts, four loads in flight per thread
__device__ __forceinline__ void csf_stage_table(double* T, const double* __restrict__ td, int ncsf, int ndp, int ts) {
    const int total = ncsf * ndp;
    for (int t0 = threadIdx.x; t0 < total; t0 += 4 * blockDim.x) {
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int t = t0 + u * int(blockDim.x);
            v[u] = t < total ? td[t] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int t = t0 + u * int(blockDim.x);
            if (t < total) {
                const int i = t / ndp;
                T[i * ts + (t - i * ndp)] = v[u];
            }
        }
    }
}

// Classes whose dense T_N fits in LDS beside the determinants of `cpb` configurations (up to 10
// open shells: 42 x 252 doubles for singlets) use dense products: no index loads, every T_N
// element and every staged determinant read from LDS, 2 x 2 outputs per thread.
constexpr size_t kCsfDenseBytes = 88u << 10;   // largest dense T_N
constexpr size_t kCsfDenseTotal = 144u << 10;  // LDS of a dense-path block
template <class TI, class TO>
__global__ __launch_bounds__(256) void k_csf_gather_dense(const TI* __restrict__ det, TO* __restrict__ csf,
                                                          const uint32_t* __restrict__ docc,
                                                          const uint32_t* __restrict__ socc,
                                                          const int64_t* __restrict__ off, int64_t nconf, int cpb,
                                                          int ncsf, int ndp, const double* __restrict__ td,
                                                          const uint32_t* __restrict__ occ, int n, int64_t nbs,
                                                          const uint32_t* __restrict__ B,
                                                          const uint16_t* __restrict__ RT) {
    __shared__ uint16_t rl[kRankTab];
    const CsfAddr addr{rl, RT != nullptr, n, nbs, B};
    csf_rank_stage(rl, RT);
    __syncthreads();
    extern __shared__ double sh[];
    const int ts = csf_tstride(ndp);
    double* T = sh;                      // [ncsf][ts]
    double* X = sh + size_t(ncsf) * ts;  // [cpb][ndp]
    const int64_t c0 = int64_t(blockIdx.x) * cpb;
    const int nc = int(min(int64_t(cpb), nconf - c0));
    csf_stage_table(T, td, ncsf, ndp, ts);
    csf_stage_dets(det, X, nc, cpb, ndp, docc + c0, socc + c0, occ, addr);
    __syncthreads();
    // 2 x 2 tiles: configurations (cc, cc + 1) x rows (i, i + hi); consecutive lanes take
    // consecutive rows, whose odd LDS stride puts them in distinct banks
    const int tc = (nc + 1) / 2, hi = (ncsf + 1) / 2;
    for (int t = threadIdx.x; t < tc * hi; t += blockDim.x) {
        const int cc = 2 * (t / hi), i = t % hi;
        const bool c2 = cc + 1 < nc, i2 = i + hi < ncsf;
        const double* x0 = X + cc * ndp;
        const double* x1 = X + (c2 ? cc + 1 : cc) * ndp;
        const double* t0 = T + i * ts;
        const double* t1 = T + (i2 ? i + hi : i) * ts;
        double a00 = 0.0, a01 = 0.0, a10 = 0.0, a11 = 0.0;
        for (int j = 0; j < ndp; ++j) {
            const double u0 = t0[j], u1 = t1[j], v0 = x0[j], v1 = x1[j];
            a00 += u0 * v0;
            a01 += u1 * v0;
            a10 += u0 * v1;
            a11 += u1 * v1;
        }
        const int64_t o0 = off[c0 + cc];
        csf[o0 + i] = TO(a00);
        if (i2) csf[o0 + i + hi] = TO(a01);
        if (c2) {
            const int64_t o1 = off[c0 + cc + 1];
            csf[o1 + i] = TO(a10);
            if (i2) csf[o1 + i + hi] = TO(a11);
        }
    }
}

// csf -> det for the same classes: det[c][j] = sum_i T[i][j] y[c][i]
template <class TO>
__global__ __launch_bounds__(256) void k_csf_scatter_dense(const double* __restrict__ csf, TO* __restrict__ det,
                                                           const uint32_t* __restrict__ docc,
                                                           const uint32_t* __restrict__ socc,
                                                           const int64_t* __restrict__ off, int64_t nconf, int cpb,
                                                           int ncsf, int ndp, const double* __restrict__ td,
                                                           const uint32_t* __restrict__ occ, int n, int64_t nbs,
                                                           const uint32_t* __restrict__ B,
                                                           const uint16_t* __restrict__ RT) {
    __shared__ uint16_t rl[kRankTab];
    const CsfAddr addr{rl, RT != nullptr, n, nbs, B};
    csf_rank_stage(rl, RT);
    __syncthreads();
    extern __shared__ double sh[];
    double* T = sh;                       // [ncsf][ndp]
    double* Y = sh + size_t(ncsf) * ndp;  // [cpb][ncsf], signed by the configuration factor
    const int64_t c0 = int64_t(blockIdx.x) * cpb;
    const int nc = int(min(int64_t(cpb), nconf - c0));
    csf_stage_table(T, td, ncsf, ndp, ndp);
    for (int t = threadIdx.x; t < cpb * ncsf; t += blockDim.x) {
        const int cc = t / ncsf, i = t - cc * ncsf;
        double v = 0.0;
        if (cc < nc) {
            v = csf[off[c0 + cc] + i];
            if (docc[c0 + cc] & kConfNeg) v = -v;
        }
        Y[t] = v;
    }
    __syncthreads();
    // 2 x 2 tiles: configurations (cc, cc + 1) x patterns (j, j + hj); consecutive lanes read
    // consecutive T_N elements
    const int tc = (nc + 1) / 2, hj = (ndp + 1) / 2;
    for (int t = threadIdx.x; t < tc * hj; t += blockDim.x) {
        const int cc = 2 * (t / hj), j = t % hj;
        const bool c2 = cc + 1 < nc, j2 = j + hj < ndp;
        const double* y0 = Y + cc * ncsf;
        const double* y1 = Y + (c2 ? cc + 1 : cc) * ncsf;
        const int jb = j2 ? j + hj : j;
        double a00 = 0.0, a01 = 0.0, a10 = 0.0, a11 = 0.0;
        for (int i = 0; i < ncsf; ++i) {
            const double u0 = T[i * ndp + j], u1 = T[i * ndp + jb], v0 = y0[i], v1 = y1[i];
            a00 += u0 * v0;
            a01 += u1 * v0;
            a10 += u0 * v1;
            a11 += u1 * v1;
        }
        for (int q = 0; q < (c2 ? 2 : 1); ++q) {
            const uint32_t dw = docc[c0 + cc + q], so = socc[c0 + cc + q];
            for (int r = 0; r < (j2 ? 2 : 1); ++r) {
                uint32_t a, b;
                csf_strings(dw & ~kConfNeg, so, occ[r == 0 ? j : jb], a, b);
                const double v = q == 0 ? (r == 0 ? a00 : a01) : (r == 0 ? a10 : a11);
                det[addr(a, b)] = TO(v);
            }
        }
    }
}

// csf -> det with the CSF coefficients of `cpb` configurations staged in LDS: work item =
// (group of kCsfCT configurations, determinant pattern j), the column j of T_N read once per group
template <class TO>
__global__ __launch_bounds__(256) void k_csf_scatter_lds(const double* __restrict__ csf, TO* __restrict__ det,
                                                         const uint32_t* __restrict__ docc,
                                                         const uint32_t* __restrict__ socc,
                                                         const int64_t* __restrict__ off, int64_t nconf, int cpb,
                                                         int ppb, int mode, int ncsf, int ndp,
                                                         const int* __restrict__ cp,
                                                         const uint32_t* __restrict__ ci, const double* __restrict__ cv,
                                                         const uint32_t* __restrict__ occ, int n, int64_t nbs,
                                                         const uint32_t* __restrict__ B,
                                                         const uint16_t* __restrict__ RT) {
    __shared__ uint16_t rl[kRankTab];
    const CsfAddr addr{rl, RT != nullptr, n, nbs, B};
    csf_rank_stage(rl, RT);
    __syncthreads();
    extern __shared__ double ys[];  // [cpb][ncsf], signed by the configuration factor
    const int64_t c0 = int64_t(blockIdx.x) * cpb;
    const int nc = int(min(int64_t(cpb), nconf - c0));
    for (int t = threadIdx.x; t < nc * ncsf; t += blockDim.x) {
        const int cc = t / ncsf, i = t - cc * ncsf;
        const double v = csf[off[c0 + cc] + i];
        ys[t] = (docc[c0 + cc] & kConfNeg) ? -v : v;
    }
    __syncthreads();
    const int ng = (nc + kCsfCT - 1) / kCsfCT;
    const int j0 = blockIdx.y * ppb, j1 = min(ndp, j0 + ppb), np = j1 - j0;
    if (mode == 2) {
        // one pattern per wave, its column's nonzeros over the lanes (hundreds at 14+ open shells)
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
        for (int j = j0 + wave; j < j1; j += nw) {
            const int e0 = cp[j], e1 = cp[j + 1];
            const uint32_t m = occ[j];
            for (int g = 0; g < ng; ++g) {
                const int cb = g * kCsfCT, kk = min(kCsfCT, nc - cb);
                const double* y0 = ys + cb * ncsf;
                double acc[kCsfCT];
#pragma unroll
                for (int k = 0; k < kCsfCT; ++k) acc[k] = 0.0;
                for (int e = e0 + lane; e < e1; e += 4 * 64) {
                    double v[4];
                    uint32_t ii[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int eu = e + 64 * u;
                        v[u] = eu < e1 ? cv[eu] : 0.0;
                        ii[u] = eu < e1 ? ci[eu] : 0u;
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u)
#pragma unroll
                        for (int k = 0; k < kCsfCT; ++k)
                            if (k < kk) acc[k] += v[u] * y0[k * ncsf + ii[u]];
                }
                double mine = 0.0;
#pragma unroll
                for (int k = 0; k < kCsfCT; ++k) {
                    const double t = csf_wave_sum(acc[k]);
                    if (lane == k) mine = t;
                }
                if (lane < kk) {
                    uint32_t a, b;
                    csf_strings(docc[c0 + cb + lane] & ~kConfNeg, socc[c0 + cb + lane], m, a, b);
                    det[addr(a, b)] = TO(mine);
                }
            }
        }
        return;
    }
    for (int t = threadIdx.x; t < ng * np; t += blockDim.x) {
        const int g = t / np, j = j0 + (t - g * np);
        const int cb = g * kCsfCT, kk = min(kCsfCT, nc - cb);
        const double* y0 = ys + cb * ncsf;
        double acc[kCsfCT];
#pragma unroll
        for (int k = 0; k < kCsfCT; ++k) acc[k] = 0.0;
        for (int e = cp[j]; e < cp[j + 1]; ++e) {
            const double v = cv[e];
            const uint32_t ii = ci[e];
#pragma unroll
            for (int k = 0; k < kCsfCT; ++k)
                if (k < kk) acc[k] += v * y0[k * ncsf + ii];
        }
        const uint32_t m = occ[j];
#pragma unroll
        for (int k = 0; k < kCsfCT; ++k)
            if (k < kk) {
                uint32_t a, b;
                csf_strings(docc[c0 + cb + k] & ~kConfNeg, socc[c0 + cb + k], m, a, b);
                det[addr(a, b)] = TO(acc[k]);
            }
    }
}

// csf -> det for classes with many configurations and long T_N columns: a wave covers cpw
// configurations (lanes) x 64 / cpw patterns; the nonzeros of a pattern's column are read by all the
// lanes of that pattern together and each lane reads its own configuration's CSF coefficients from
// LDS (row stride ncsf | 1, so the lanes' rows fall in distinct banks)
template <class TO>
__global__ __launch_bounds__(256) void k_csf_scatter_cfg(const double* __restrict__ csf, TO* __restrict__ det,
                                                         const uint32_t* __restrict__ docc,
                                                         const uint32_t* __restrict__ socc,
                                                         const int64_t* __restrict__ off, int64_t nconf, int cpw,
                                                         int ppb, int ncsf, int ndp, const int* __restrict__ cp,
                                                         const uint32_t* __restrict__ ci, const double* __restrict__ cv,
                                                         const uint32_t* __restrict__ occ, int n, int64_t nbs,
                                                         const uint32_t* __restrict__ B,
                                                         const uint16_t* __restrict__ RT) {
    __shared__ uint16_t rl[kRankTab];
    const CsfAddr addr{rl, RT != nullptr, n, nbs, B};
    csf_rank_stage(rl, RT);
    extern __shared__ double ys[];  // [cpw][ncsf | 1], signed by the configuration factor
    const int ys_st = ncsf | 1;
    const int64_t c0 = int64_t(blockIdx.x) * cpw;
    const int nc = int(min(int64_t(cpw), nconf - c0));
    for (int t = threadIdx.x; t < nc * ncsf; t += blockDim.x) {
        const int cc = t / ncsf, i = t - cc * ncsf;
        const double v = csf[off[c0 + cc] + i];
        ys[cc * ys_st + i] = (docc[c0 + cc] & kConfNeg) ? -v : v;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int cc = lane % cpw, sub = lane / cpw, pw = 64 / cpw;  // patterns per wave
    const bool live = cc < nc;
    const uint32_t dw = live ? docc[c0 + cc] : 0u, so = live ? socc[c0 + cc] : 0u;
    const double* y = ys + (live ? cc : 0) * ys_st;
    const int j0 = blockIdx.y * ppb, j1 = min(ndp, j0 + ppb);
    for (int jb = j0 + wave * pw; jb < j1; jb += 4 * pw) {
        const int j = jb + sub;
        const bool ok = j < j1;
        const int e0 = ok ? cp[j] : 0, e1 = ok ? cp[j + 1] : 0;
        double acc = 0.0;
        int e = e0;
        for (; e + 4 <= e1; e += 4) {
            double v[4];
            uint32_t ii[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                v[u] = cv[e + u];
                ii[u] = ci[e + u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) acc += v[u] * y[ii[u]];
        }
        for (; e < e1; ++e) acc += cv[e] * y[ci[e]];
        if (live && ok) {
            uint32_t a, b;
            csf_strings(dw & ~kConfNeg, so, occ[j], a, b);
            det[addr(a, b)] = TO(acc);
        }
    }
}

// csf -> det: thread per (configuration, determinant pattern); every determinant of the sector
// belongs to exactly one configuration, so each is written once
template <class TO>
__global__ void k_csf_scatter(const double* __restrict__ csf, TO* __restrict__ det, const uint32_t* __restrict__ docc,
                              const uint32_t* __restrict__ socc, const int64_t* __restrict__ off, int64_t nconf,
                              int ndp, const int* __restrict__ cp, const uint32_t* __restrict__ ci,
                              const double* __restrict__ cv, const uint32_t* __restrict__ occ, int n, int64_t nbs,
                              const uint32_t* __restrict__ B, const uint16_t* __restrict__ RT) {
    __shared__ uint16_t rl[kRankTab];
    const CsfAddr addr{rl, RT != nullptr, n, nbs, B};
    csf_rank_stage(rl, RT);
    __syncthreads();
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
    det[addr(a, b)] = TO((dw & kConfNeg) ? -acc : acc);
}

// FCISolver::form_Hdiag_csf without the full preconditioner (fci_solver_initial_guess.cc:134-141):
// sum_j T[i][j]^2 H_jj
__global__ void k_csf_hdiag(const double* __restrict__ hd, double* __restrict__ out, const uint32_t* __restrict__ docc,
                            const uint32_t* __restrict__ socc, const int64_t* __restrict__ off, int64_t nconf,
                            int ncsf, const int* __restrict__ rp, const uint32_t* __restrict__ rj,
                            const double* __restrict__ rv, const uint32_t* __restrict__ occ, int n, int64_t nbs,
                            const uint32_t* __restrict__ B, const uint16_t* __restrict__ RT) {
    __shared__ uint16_t rl[kRankTab];
    const CsfAddr addr{rl, RT != nullptr, n, nbs, B};
    csf_rank_stage(rl, RT);
    __syncthreads();
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nconf * ncsf) return;
    const int64_t c = t / ncsf;
    const int i = int(t % ncsf);
    const uint32_t d = docc[c] & ~kConfNeg, so = socc[c];
    double acc = 0.0;
    for (int e = rp[i]; e < rp[i + 1]; ++e) {
        uint32_t a, b;
        csf_strings(d, so, occ[rj[e]], a, b);
        acc += rv[e] * rv[e] * hd[addr(a, b)];
    }
    out[off[c] + i] = acc;
}

inline unsigned grid_of(int64_t work) { return unsigned((work + 255) / 256); }
// largest staged determinant row of one configuration (12,870 doubles for 16 open shells)
constexpr size_t kCsfLdsBytes = 128u << 10;

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

CsfSpace::CsfSpace(int n, int na, int nb, int twoS, const std::vector<int>* orbsym, int irrep, const GasSpec* gas,
                   const uint32_t* d_binom, int64_t nbs)
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
                if (gas) {  // electrons per GAS space: 2 per closed shell, 1 per open shell
                    bool ok = true;
                    for (size_t g = 0; g < gas->orbs.size() && ok; ++g) {
                        const int e = 2 * __builtin_popcount(dm & gas->orbs[g]) + __builtin_popcount(so & gas->orbs[g]);
                        ok = e >= gas->gmin[g] && e <= gas->gmax[g];
                    }
                    if (!ok) continue;
                }
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
    if (n <= 16) {
        // H[v]: bits of v at orbitals 8..15, ordinals from the top; L[c][v]: bits of v at orbitals
        // 0..7 below c electrons of the high byte (entries of byte values no string has are unused)
        std::vector<uint16_t> tab(kRankTab, 0);
        auto part = [&](uint32_t v, int base, int c) {
            uint64_t r = 0;
            for (int i = 7; i >= 0; --i)
                if ((v >> i) & 1u) {
                    const int orb = base + i;
                    if (orb >= n) return uint16_t(0);
                    r += binom_host(n - 1 - orb, 1 + c + __builtin_popcount(v >> (i + 1)));
                }
            return uint16_t(r);
        };
        for (uint32_t v = 0; v < 256; ++v) {
            tab[v] = part(v, 8, 0);
            for (int c = 0; c <= 8; ++c) tab[256 + c * 256 + v] = part(v, 0, c);
        }
        rtab_.alloc(tab.size());
        rtab_.upload(tab.data(), tab.size());
    }
    for (auto& C : cls_) {
        C.nconf = int64_t(C.h_docc.size());
        ndet_cov_ += C.nconf * C.ndp;
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
        if (size_t(C.ncsf) * C.ndp * sizeof(double) <= kCsfDenseBytes) {
            std::vector<double> dense(size_t(C.ncsf) * C.ndp, 0.0);
            for (int i = 0; i < C.ncsf; ++i)
                for (int e = C.h_rp[i]; e < C.h_rp[i + 1]; ++e) dense[size_t(i) * C.ndp + C.h_rj[e]] = C.h_rv[e];
            C.td.alloc(dense.size());
            C.td.upload(dense.data(), dense.size());
        }
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
        const size_t row_bytes = size_t(C.ndp) * sizeof(double);
        const size_t tbytes = size_t(C.ncsf) * csf_tstride(C.ndp) * sizeof(double);
        if (C.td.p && tbytes + 8 * row_bytes <= kCsfDenseTotal) {
            const int cpb = int(std::min<size_t>(64, (kCsfDenseTotal - tbytes) / row_bytes));
            k_csf_gather_dense<TI, TO><<<unsigned((C.nconf + cpb - 1) / cpb), 256, tbytes + size_t(cpb) * row_bytes, st>>>(
                det, csf, C.docc.p, C.socc.p, C.off.p, C.nconf, cpb, C.ncsf, C.ndp, C.td.p, C.occ.p, n_, nbs_, binom_, rtab_.p);
        } else if (row_bytes <= kCsfLdsBytes) {
            // up to as many configurations per block as fit in 64 KB of LDS (at least one), fewer
            // when that leaves under 512 blocks; rows split over blocks so that classes with few
            // configurations still fill the chip (each split stages the determinants again)
            const int64_t fit = int64_t(std::max<size_t>(1, std::min<size_t>(64, (64u << 10) / row_bytes)));
            const int cpb = int(std::max<int64_t>(1, std::min<int64_t>(fit, C.nconf / 512)));
            const int64_t nxb = (C.nconf + cpb - 1) / cpb;
            const double per_row = double(C.h_rj.size()) / C.ncsf;
            const int mode = per_row >= 32.0 ? 2 : (C.ncsf < 64 ? 0 : 1);
            int rpb = mode == 0 ? C.ncsf : 256;
            if (mode == 2) {
                const int64_t nyw = std::max<int64_t>(1, std::min<int64_t>((1024 + nxb - 1) / nxb, (C.ncsf + 7) / 8));
                rpb = int((C.ncsf + nyw - 1) / nyw);
            }
            const unsigned nyb = unsigned((C.ncsf + rpb - 1) / rpb);
            k_csf_gather_lds<TI, TO><<<dim3(unsigned(nxb), nyb), 256, size_t(cpb) * row_bytes, st>>>(
                det, csf, C.docc.p, C.socc.p, C.off.p, C.nconf, cpb, rpb, mode, C.ncsf, C.ndp, C.rp.p, C.rj.p, C.rv.p,
                C.occ.p, n_, nbs_, binom_, rtab_.p);
        } else {
            k_csf_gather<TI, TO><<<grid_of(C.nconf * C.ncsf), 256, 0, st>>>(
                det, csf, C.docc.p, C.socc.p, C.off.p, C.nconf, C.ncsf, C.rp.p, C.rj.p, C.rv.p, C.occ.p, n_, nbs_,
                binom_, rtab_.p);
        }
    }
    HIPCHK(hipGetLastError());
}

template <class TO> void CsfSpace::to_det(const double* csf, TO* det, int64_t ndet, hipStream_t st) const {
    // every determinant of the configurations is written once; the others (other irreps, other
    // GAS occupations, configurations with too few open shells) are zero
    if (ndet_cov_ != ndet) HIPCHK(hipMemsetAsync(det, 0, size_t(ndet) * sizeof(TO), st));
    for (const auto& C : cls_) {
        if (C.nconf == 0) continue;
        const size_t tbytes = size_t(C.ncsf) * C.ndp * sizeof(double);
        const size_t ybytes = size_t(std::max(1, C.ncsf)) * sizeof(double);
        // configurations per block: enough (config group, pattern) items for 256 threads, their
        // CSF coefficients within 32 KB of LDS; patterns split over blocks when there are few
        const int want = int(std::max<int64_t>(1, (256 * kCsfCT + C.ndp - 1) / C.ndp));
        const int cpb = std::max(1, std::min<int>(want, int((32u << 10) / ybytes)));
        if (C.td.p && tbytes + 8 * ybytes <= kCsfDenseTotal) {
            const int cpd = int(std::min<size_t>(64, (kCsfDenseTotal - tbytes) / ybytes));
            k_csf_scatter_dense<TO><<<unsigned((C.nconf + cpd - 1) / cpd), 256, tbytes + size_t(cpd) * ybytes, st>>>(
                csf, det, C.docc.p, C.socc.p, C.off.p, C.nconf, cpd, C.ncsf, C.ndp, C.td.p, C.occ.p, n_, nbs_, binom_, rtab_.p);
        } else if (C.ncsf > 0 && ybytes <= (32u << 10)) {
            const double per_col = double(C.h_rj.size()) / C.ndp;  // nonzeros per T_N column
            // lanes over configurations: as many (a power of two up to 64) as fit in 72 KB
            int cpw = 64;
            while (cpw > 8 && size_t(cpw) * (C.ncsf | 1) * sizeof(double) > (72u << 10)) cpw /= 2;
            const bool fits = size_t(cpw) * (C.ncsf | 1) * sizeof(double) <= (72u << 10);
            if (per_col >= 16.0 && fits && C.nconf >= 8) {
                const int64_t nxb = (C.nconf + cpw - 1) / cpw;
                const int step = 4 * (64 / cpw);  // patterns per pass of a block
                const int64_t nyw = std::max<int64_t>(1, std::min<int64_t>((1024 + nxb - 1) / nxb, C.ndp / step));
                const int ppb = int(((C.ndp + nyw - 1) / nyw + step - 1) / step * step);
                k_csf_scatter_cfg<TO><<<dim3(unsigned(nxb), unsigned((C.ndp + ppb - 1) / ppb)), 256,
                                        size_t(cpw) * (C.ncsf | 1) * sizeof(double), st>>>(
                    csf, det, C.docc.p, C.socc.p, C.off.p, C.nconf, cpw, ppb, C.ncsf, C.ndp, C.cp.p, C.ci.p, C.cv.p,
                    C.occ.p, n_, nbs_, binom_, rtab_.p);
            } else if (per_col >= 16.0) {
                // few configurations: one pattern per wave, lanes over its column
                const int cpw2 = int(std::max<int64_t>(1, std::min<int64_t>(C.nconf, std::min<int>(8, cpb))));
                const int64_t nxb = (C.nconf + cpw2 - 1) / cpw2;
                const int ppb = 16;
                k_csf_scatter_lds<TO><<<dim3(unsigned(nxb), unsigned((C.ndp + ppb - 1) / ppb)), 256,
                                        size_t(cpw2) * ybytes, st>>>(csf, det, C.docc.p, C.socc.p, C.off.p, C.nconf,
                                                                     cpw2, ppb, 2, C.ncsf, C.ndp, C.cp.p, C.ci.p,
                                                                     C.cv.p, C.occ.p, n_, nbs_, binom_, rtab_.p);
            } else {
                const int64_t nxb = (C.nconf + cpb - 1) / cpb;
                const int ppb = nxb >= 1024 ? C.ndp : 256;
                k_csf_scatter_lds<TO><<<dim3(unsigned(nxb), unsigned((C.ndp + ppb - 1) / ppb)), 256,
                                        size_t(cpb) * ybytes, st>>>(csf, det, C.docc.p, C.socc.p, C.off.p, C.nconf,
                                                                    cpb, ppb, 0, C.ncsf, C.ndp, C.cp.p, C.ci.p,
                                                                    C.cv.p, C.occ.p, n_, nbs_, binom_, rtab_.p);
            }
        } else {
            k_csf_scatter<TO><<<grid_of(C.nconf * C.ndp), 256, 0, st>>>(
                csf, det, C.docc.p, C.socc.p, C.off.p, C.nconf, C.ndp, C.cp.p, C.ci.p, C.cv.p, C.occ.p, n_, nbs_,
                binom_, rtab_.p);
        }
    }
    HIPCHK(hipGetLastError());
}

void CsfSpace::hdiag(const double* hdet, double* hcsf, hipStream_t st) const {
    for (const auto& C : cls_) {
        if (C.nconf == 0 || C.ncsf == 0) continue;
        k_csf_hdiag<<<grid_of(C.nconf * C.ncsf), 256, 0, st>>>(hdet, hcsf, C.docc.p, C.socc.p, C.off.p, C.nconf,
                                                              C.ncsf, C.rp.p, C.rj.p, C.rv.p, C.occ.p, n_, nbs_,
                                                              binom_, rtab_.p);
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
