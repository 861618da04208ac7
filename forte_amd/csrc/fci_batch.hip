// fci_batch.hip -- many independent FCI problems on one GPU (CASSCF micro-iterations, geometry
// scans): every problem is Forte's FCISolver::compute_energy (forte/fci/fci_solver.cc:177-314) with
// its Davidson-Liu loop (forte/helpers/davidson_liu_solver.cc:180-273), FP64, diagonal
// preconditioner.
//
// Problems of one active-space shape (n, na, nb) share the string tables and advance together.
// One Davidson-Liu iteration of the whole class is a fixed sequence of batched kernels: the sigma
// terms run once per term over all (problem, new vector) items (batch_sigma_f64), and the
// subspace algebra (b.sigma overlaps, the Rayleigh-Ritz eigenproblem by a one-wave Jacobi per
// problem, residuals and the preconditioner, the projection of other-multiplicity roots, collapse
// and Gram-Schmidt with the reference's discard / orthogonality thresholds, convergence) reads and
// writes per-problem state in device memory, so an iteration needs no host decision. That
// sequence is captured once in a hipGraph and replayed; the host only polls, every few
// iterations, how many problems are done. Converged problems freeze: every kernel skips them.
// The streaming kernels' chunk length and the sigma sub-batch size are chosen per class by timing
// the candidates on the class's own first iterations (the autotuner below).
#include "../../include/forte_fci.h"
#include "fci_dl.h"
#include "fci_internal.h"

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cmath>
#include <atomic>
#include <cstring>
#include <exception>
#include <map>
#include <system_error>
#include <thread>

namespace fci {

void check_device_public();

constexpr int BMS = 32;         // max subspace size (subspace_per_root * nroot)
constexpr int BNR = 4;          // max roots
// Gram-Schmidt passes per added vector inside one iteration: a correction needs a measuring pass,
// one or two subtractions and the scaling pass; a collapsed vector is orthonormal up to rounding
// (a measuring and a scaling pass); a vector not accepted within its passes is discarded
constexpr int BGS_PASSES = 6;   // guesses and final Ritz vectors (outside the graph)
constexpr int BGS_PASSES_CORR = 5, BGS_PASSES_COLL = 4;
constexpr int BBS = 256;        // threads of the streaming kernels
constexpr int BPS = BMS * BMS;  // partial-sum stride per block (largest: the overlap matrix)
constexpr int BNPB = 128;       // at most this many blocks per problem in a streaming kernel

struct BState {
    int bs, ss, gv;      // basis size, vectors with sigma, overlap entries valid
    int done, conv, status, r_ok;
    int iters, nsigma;
    int ntoadd, added, collapse, cs, nmiss;
    // Gram-Schmidt of one source vector into row gs_row of the basis
    int gs_active, gs_scaled, gs_ok, gs_row;
    double gs_a;
    const double* gs_src;
    double lam[BNR], lam_old[BNR], res[BNR], rnorm[BNR];
    double G[BMS * BMS];      // b_i . sigma_j
    double alpha[BMS * BMS];  // Ritz coefficients alpha[i * BMS + k]
    double gs_d[BMS];
    unsigned long long rng;
};

struct BDev {
    int P, NR, MS, CS, NSL;
    int64_t nd;
    int nch;     // chunks of chl elements per vector
    int npb;     // blocks per problem in the streaming kernels (each loops over its chunks)
    int64_t chl;
    BState* st;
    double *b, *sg, *r, *tmp, *tms;  // [P][MS][nd], [P][MS][nd], [P][NR][nd], [P][TM][nd], [P][CS][nd]
    int TM;
    const double* hd;                // [P][nd]
    double* part;                    // [P][npb][BPS]: one partial per block
    const int* pvo;                  // bad roots of problem p: vectors [pvo[p], pvo[p+1])
    const int64_t* pvoff;            // entries of bad vector v: [pvoff[v], pvoff[v+1])
    const int64_t* pidx;
    const double* pval;
    double e_conv, r_conv, discard, orth;
    int maxiter;
};

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}
// sum over a 256-thread block, result valid in thread 0
__device__ __forceinline__ double block_sum(double v, double* red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < BBS / 64; ++i) s += red[i];
    __syncthreads();
    return s;
}

__device__ __forceinline__ double* vec(double* base, int64_t nd, int64_t row) { return base + row * nd; }

// ---- setup --------------------------------------------------------------------------------------
__global__ void k_b_init(BDev d, int ng) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.P) return;
    BState& s = d.st[p];
    s.bs = 0; s.ss = 0; s.gv = 0;
    s.done = 0; s.conv = 0; s.status = 0; s.r_ok = 0;
    s.iters = 0; s.nsigma = 0;
    s.ntoadd = ng; s.added = 0; s.collapse = 0; s.cs = 0; s.nmiss = 0;
    s.gs_active = 0; s.gs_scaled = 0; s.gs_ok = 0; s.gs_row = 0; s.gs_a = 1.0; s.gs_src = nullptr;
    for (int k = 0; k < BNR; ++k) { s.lam[k] = 0.0; s.lam_old[k] = 0.0; s.res[k] = 0.0; s.rnorm[k] = 1.0; }
    s.rng = 0x5EEDULL + 0x9E3779B97F4A7C15ULL * (unsigned long long)(p + 1);
}

// guess vectors (sparse) into tmp: entries [goff[p*ng+g], goff[p*ng+g+1]) of (gidx, gval)
__global__ void k_b_scatter_guess(BDev d, int ng, const int64_t* goff, const int64_t* gidx, const double* gval) {
    const int p = blockIdx.y, g = blockIdx.z;
    const int64_t v = int64_t(p) * ng + g;
    double* t = vec(d.tmp, d.nd, int64_t(p) * d.TM + g);
    for (int64_t e = goff[v] + blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < goff[v + 1];
         e += int64_t(gridDim.x) * blockDim.x)
        t[gidx[e]] = gval[e];  // set_vector: later duplicates overwrite (entries are unique here)
}

__global__ void k_b_zero(double* x, int64_t n) {
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        x[i] = 0.0;
}

// ---- sigma bookkeeping --------------------------------------------------------------------------
// item (p, s): vector ss + s of problem p needs its sigma when it exists and p is not done
__global__ void k_b_sig_tab(BDev d, const double** c_tab, double** s_tab) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.P * d.NSL) return;
    const int p = i / d.NSL, sl = i % d.NSL;
    const BState& s = d.st[p];
    const int j = s.ss + sl;
    const bool on = !s.done && j < s.bs;
    c_tab[i] = on ? vec(d.b, d.nd, int64_t(p) * d.MS + j) : nullptr;
    s_tab[i] = on ? vec(d.sg, d.nd, int64_t(p) * d.MS + j) : nullptr;
}

__global__ void k_b_after_sigma(BDev d) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.P) return;
    BState& s = d.st[p];
    if (s.done) return;
    s.iters += 1;
    s.nsigma += s.bs - s.ss;
    s.ss = s.bs;
}

// ---- Rayleigh-Ritz ------------------------------------------------------------------------------
// partial sums of G[i][j] = b_i . sigma_j for the entries not yet known (max(i, j) >= gv): the
// chunk's elements are staged 32 at a time for every basis and sigma vector in LDS, so each
// element is read from memory once, and each thread accumulates up to 4 of the pairs
constexpr int GD_TE = 32;
__global__ __launch_bounds__(BBS) void k_b_gdots(BDev d) {
    __shared__ double Bt[BMS][GD_TE + 1], St[BMS][GD_TE + 1];
    const int p = blockIdx.y;
    const BState& s = d.st[p];
    if (s.done) return;
    const int bs = s.bs, gv = s.gv;
    const int npair = bs * bs - gv * gv;
    const double* B = d.b + int64_t(p) * d.MS * d.nd;
    const double* S = d.sg + int64_t(p) * d.MS * d.nd;
    // pair index t -> (i, j): the rows i >= gv (all columns), then rows i < gv with columns j >= gv
    auto pair_of = [&](int t, int& i, int& j) {
        const int nhi = (bs - gv) * bs;
        if (t < nhi) {
            i = gv + t / bs;
            j = t % bs;
        } else {
            t -= nhi;
            i = t / (bs - gv);
            j = gv + t % (bs - gv);
        }
    };
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    int pi[4], pj[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int t = threadIdx.x + u * BBS;
        pi[u] = pj[u] = 0;
        if (t < npair) pair_of(t, pi[u], pj[u]);
    }
    for (int c = blockIdx.x; c < d.nch; c += gridDim.x)
    for (int64_t t0 = int64_t(c) * d.chl, e1 = min(t0 + d.chl, d.nd); t0 < e1; t0 += GD_TE) {
        for (int x = threadIdx.x; x < bs * GD_TE; x += BBS) {
            const int v = x / GD_TE, e = x % GD_TE;
            const int64_t el = t0 + e;
            const bool ok = el < e1;
            Bt[v][e] = ok ? B[int64_t(v) * d.nd + el] : 0.0;
            St[v][e] = ok ? S[int64_t(v) * d.nd + el] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (threadIdx.x + u * BBS >= npair) break;
            double a = acc[u];
#pragma unroll 8
            for (int e = 0; e < GD_TE; ++e) a += Bt[pi[u]][e] * St[pj[u]][e];
            acc[u] = a;
        }
        __syncthreads();
    }
    double* out = d.part + (int64_t(p) * d.npb + blockIdx.x) * BPS;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int t = threadIdx.x + u * BBS;
        if (t < npair) out[pi[u] * BMS + pj[u]] = acc[u];
    }
}

// One wave per problem: complete G, then Jacobi on sym(G) (the host solver's rotation formula,
// sweep rule and ascending stable order, with the rotations of a sweep applied in parallel
// rounds), lambda and alpha.
__global__ __launch_bounds__(64) void k_b_eigh(BDev d) {
    __shared__ double A[BMS][BMS + 1], V[BMS][BMS + 1];
    __shared__ double ev[BMS];
    __shared__ int order[BMS];
    const int p = blockIdx.x, l = threadIdx.x;
    BState& s = d.st[p];
    if (s.done) return;
    const int n = s.bs, gv = s.gv;
    // G entries from the chunk partials, summed in chunk order
    for (int idx = l; idx < n * n; idx += 64) {
        const int i = idx / n, j = idx % n;
        if (i < gv && j < gv) continue;
        double acc = 0.0;
        for (int c = 0; c < d.npb; ++c) acc += d.part[(int64_t(p) * d.npb + c) * BPS + i * BMS + j];
        s.G[i * BMS + j] = acc;
    }
    __syncthreads();
    if (l == 0) s.gv = n;
    for (int idx = l; idx < n * n; idx += 64) {
        const int i = idx / n, j = idx % n;
        A[i][j] = 0.5 * (s.G[i * BMS + j] + s.G[j * BMS + i]);
        V[i][j] = i == j ? 1.0 : 0.0;
    }
    __syncthreads();
    // parallel-order Jacobi: each round applies m/2 disjoint rotations at once (round-robin
    // pairing of m = n rounded up to even indices; pairs with the padding index are skipped), so a
    // sweep is m - 1 rounds instead of n(n-1)/2 dependent rotations
    __shared__ double rc[BMS / 2 + 1], rs[BMS / 2 + 1];
    __shared__ int rp[BMS / 2 + 1], rq[BMS / 2 + 1];
    const int m = n + (n & 1);
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int idx = l; idx < n * n; idx += 64) {
            const int i = idx / n, j = idx % n;
            const double a = A[i][j];
            tot += a * a;
            if (i != j) off += a * a;
        }
        off = __shfl(wave_sum(off), 0, 64);
        tot = __shfl(wave_sum(tot), 0, 64);
        if (off == 0.0 || off <= 1e-30 * (tot > 0 ? tot : 1.0)) break;
        for (int r = 0; r < m - 1; ++r) {
            if (l < m / 2) {
                int p0, q0;
                if (l == 0) {
                    p0 = r;
                    q0 = m - 1;
                } else {
                    p0 = (r + l) % (m - 1);
                    q0 = (r - l + m - 1) % (m - 1);
                }
                const int pp = min(p0, q0), q = max(p0, q0);
                double cc = 1.0, sn = 0.0;
                if (q < n) {
                    const double apq = A[pp][q];
                    if (fabs(apq) >= 1e-300) {
                        const double theta = (A[q][q] - A[pp][pp]) / (2.0 * apq);
                        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                        cc = 1.0 / sqrt(t * t + 1.0);
                        sn = t * cc;
                    }
                }
                rc[l] = cc;
                rs[l] = sn;
                rp[l] = pp;
                rq[l] = q < n ? q : pp;  // a pair with the padding index is the identity
            }
            __syncthreads();
            if (l < n) {  // columns p, q of row l, and of V
                for (int k = 0; k < m / 2; ++k) {
                    const int pp = rp[k], q = rq[k];
                    if (pp == q) continue;
                    const double cc = rc[k], sn = rs[k];
                    const double akp = A[l][pp], akq = A[l][q];
                    A[l][pp] = cc * akp - sn * akq;
                    A[l][q] = sn * akp + cc * akq;
                    const double vkp = V[l][pp], vkq = V[l][q];
                    V[l][pp] = cc * vkp - sn * vkq;
                    V[l][q] = sn * vkp + cc * vkq;
                }
            }
            __syncthreads();
            if (l < n) {  // rows p, q of column l
                for (int k = 0; k < m / 2; ++k) {
                    const int pp = rp[k], q = rq[k];
                    if (pp == q) continue;
                    const double cc = rc[k], sn = rs[k];
                    const double apk = A[pp][l], aqk = A[q][l];
                    A[pp][l] = cc * apk - sn * aqk;
                    A[q][l] = sn * apk + cc * aqk;
                }
            }
            __syncthreads();
        }
    }
    if (l == 0) {
        for (int i = 0; i < n; ++i) {
            order[i] = i;
            ev[i] = A[i][i];
        }
        for (int i = 1; i < n; ++i) {  // stable insertion sort, ascending
            const int oi = order[i];
            int j = i - 1;
            while (j >= 0 && ev[order[j]] > ev[oi]) {
                order[j + 1] = order[j];
                --j;
            }
            order[j + 1] = oi;
        }
    }
    __syncthreads();
    for (int idx = l; idx < BMS * BMS; idx += 64) s.alpha[idx] = 0.0;
    __syncthreads();
    for (int idx = l; idx < n * n; idx += 64) {
        const int i = idx / n, k = idx % n;
        s.alpha[i * BMS + k] = V[i][order[k]];
    }
    if (l < d.NR) s.lam[l] = l < n ? ev[order[l]] : 0.0;
}

// r_k = sum_j alpha_jk (sigma_j - lambda_k b_j); partial |r_k|^2; then the diagonal
// preconditioner in place (davidson_liu_solver.cc:412-448)
__global__ __launch_bounds__(BBS) void k_b_residual(BDev d) {
    __shared__ double red[BBS / 64];
    __shared__ double al[BMS * BNR];
    const int p = blockIdx.y;
    const BState& s = d.st[p];
    if (s.done) return;
    const int bs = s.bs, NR = d.NR;
    for (int i = threadIdx.x; i < bs * NR; i += BBS) al[i] = s.alpha[(i / NR) * BMS + i % NR];
    __syncthreads();
    double lam[BNR];
    for (int k = 0; k < BNR; ++k) lam[k] = k < NR ? s.lam[k] : 0.0;
    const double* B = d.b + int64_t(p) * d.MS * d.nd;
    const double* S = d.sg + int64_t(p) * d.MS * d.nd;
    const double* H = d.hd + int64_t(p) * d.nd;
    double* R = d.r + int64_t(p) * NR * d.nd;
    double nrm[BNR] = {0.0, 0.0, 0.0, 0.0};
    for (int c = blockIdx.x; c < d.nch; c += gridDim.x)
    for (int64_t e = int64_t(c) * d.chl + threadIdx.x, e1 = min(int64_t(c) * d.chl + d.chl, d.nd); e < e1; e += BBS) {
        double sa[BNR] = {0.0, 0.0, 0.0, 0.0}, ba[BNR] = {0.0, 0.0, 0.0, 0.0};
        for (int j = 0; j < bs; ++j) {
            const double sj = S[int64_t(j) * d.nd + e], bj = B[int64_t(j) * d.nd + e];
#pragma unroll
            for (int k = 0; k < BNR; ++k)
                if (k < NR) {
                    sa[k] += al[j * NR + k] * sj;
                    ba[k] += al[j * NR + k] * bj;
                }
        }
        const double h = H[e];
#pragma unroll
        for (int k = 0; k < BNR; ++k)
            if (k < NR) {
                const double rk = sa[k] - lam[k] * ba[k];
                nrm[k] += rk * rk;
                const double den = lam[k] - h;
                R[int64_t(k) * d.nd + e] = fabs(den) > 1.0e-6 ? rk / den : 0.0;
            }
    }
    double* out = d.part + (int64_t(p) * d.npb + blockIdx.x) * BPS;
    for (int k = 0; k < NR; ++k) {
        const double t = block_sum(nrm[k], red);
        if (threadIdx.x == 0) out[k] = t;
    }
}

// residual norms, convergence (davidson_liu_solver.cc:467-486), collapse decision
__global__ void k_b_converge(BDev d) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.P) return;
    BState& s = d.st[p];
    if (s.done) return;
    int ne = 0, nr = 0;
    for (int k = 0; k < d.NR; ++k) {
        double acc = 0.0;
        for (int c = 0; c < d.npb; ++c) acc += d.part[(int64_t(p) * d.npb + c) * BPS + k];
        s.res[k] = sqrt(acc);
        ne += fabs(s.lam[k] - s.lam_old[k]) < d.e_conv;
        nr += s.res[k] < d.r_conv;
    }
    s.r_ok = nr == d.NR;
    if ((ne == d.NR && s.r_ok) || int64_t(s.bs) == d.nd) {
        s.done = 1;
        s.conv = 1;
        return;
    }
    if (s.iters >= d.maxiter) {
        s.done = 1;
        s.conv = 0;
        return;
    }
    s.collapse = s.bs + d.NR > d.MS;
    s.cs = min(max(d.CS, d.NR), s.bs);
    s.ntoadd = min(d.NR, d.MS - (s.collapse ? s.cs : s.bs));
    s.added = 0;
}

// v <- v - sum_b (v . pb) pb over the problem's other-multiplicity roots (sparse), for rows
// [0, nvec) of the given per-problem vector block; each vector is one block
__global__ __launch_bounds__(BBS) void k_b_project(BDev d, double* base, int per_prob, int which) {
    __shared__ double red[BBS / 64];
    __shared__ double dot;
    const int p = blockIdx.x, k = blockIdx.y;
    const BState& s = d.st[p];
    if (s.done) return;
    // which: 0 = corrections (all NR), 1 = random fill (the first nmiss, at most NR)
    if (which == 1 && k >= min(s.nmiss, d.NR)) return;
    double* v = base + (int64_t(p) * per_prob + k) * d.nd;
    for (int b = d.pvo[p]; b < d.pvo[p + 1]; ++b) {
        double acc = 0.0;
        for (int64_t e = d.pvoff[b] + threadIdx.x; e < d.pvoff[b + 1]; e += BBS) acc += v[d.pidx[e]] * d.pval[e];
        acc = block_sum(acc, red);
        if (threadIdx.x == 0) dot = acc;
        __syncthreads();
        const double dd = dot;
        for (int64_t e = d.pvoff[b] + threadIdx.x; e < d.pvoff[b + 1]; e += BBS) v[d.pidx[e]] -= dd * d.pval[e];
        __syncthreads();
    }
}

__global__ __launch_bounds__(BBS) void k_b_rnorm_part(BDev d) {
    __shared__ double red[BBS / 64];
    const int p = blockIdx.y;
    if (d.st[p].done) return;
    double* out = d.part + (int64_t(p) * d.npb + blockIdx.x) * BPS;
    for (int k = 0; k < d.NR; ++k) {
        const double* R = d.r + (int64_t(p) * d.NR + k) * d.nd;
        double acc = 0.0;
        for (int c = blockIdx.x; c < d.nch; c += gridDim.x)
            for (int64_t e = int64_t(c) * d.chl + threadIdx.x, e1 = min(int64_t(c) * d.chl + d.chl, d.nd); e < e1;
                 e += BBS)
                acc += R[e] * R[e];
        acc = block_sum(acc, red);
        if (threadIdx.x == 0) out[k] = acc;
    }
}

__global__ void k_b_rnorm(BDev d) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.P) return;
    BState& s = d.st[p];
    if (s.done) return;
    for (int k = 0; k < d.NR; ++k) {
        double acc = 0.0;
        for (int c = 0; c < d.npb; ++c) acc += d.part[(int64_t(p) * d.npb + c) * BPS + k];
        s.rnorm[k] = sqrt(acc);
    }
}

// ---- collapse (davidson_liu_solver.cc:538-588) and the final Ritz vectors -------------------------
// out_k = sum_j alpha_jk X_j for k < kout (kout: collapse size, or nroot when final)
__global__ __launch_bounds__(BBS) void k_b_rotate(BDev d, const double* X, int xper, double* out, int oper,
                                                  int final_) {
    __shared__ double al[BMS * BMS];
    const int p = blockIdx.y;
    const BState& s = d.st[p];
    if (!final_ && (s.done || !s.collapse)) return;
    const int bs = s.bs, kout = final_ ? d.NR : s.cs;
    for (int i = threadIdx.x; i < bs * kout; i += BBS) al[i] = s.alpha[(i / kout) * BMS + i % kout];
    __syncthreads();
    const double* Xp = X + int64_t(p) * xper * d.nd;
    double* Op = out + int64_t(p) * oper * d.nd;
    for (int c = blockIdx.x; c < d.nch; c += gridDim.x)
    for (int64_t e = int64_t(c) * d.chl + threadIdx.x, e1 = min(int64_t(c) * d.chl + d.chl, d.nd); e < e1; e += BBS) {
        double acc[BMS];
#pragma unroll
        for (int k = 0; k < BMS; ++k) acc[k] = 0.0;
        for (int j = 0; j < bs; ++j) {
            const double x = Xp[int64_t(j) * d.nd + e];
#pragma unroll
            for (int k = 0; k < BMS; ++k)
                if (k < kout) acc[k] += al[j * kout + k] * x;
        }
#pragma unroll
        for (int k = 0; k < BMS; ++k)
            if (k < kout) Op[int64_t(k) * d.nd + e] = acc[k];
    }
}

__global__ void k_b_collapse_begin(BDev d) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.P) return;
    BState& s = d.st[p];
    if (s.done || !s.collapse) return;
    s.added = 0;  // rows of the collapsed basis accepted so far
}

__global__ __launch_bounds__(BBS) void k_b_copy_sigma(BDev d) {
    const int p = blockIdx.y;
    const BState& s = d.st[p];
    if (s.done || !s.collapse) return;
    for (int k = 0; k < s.cs; ++k) {
        const double* src = d.tms + (int64_t(p) * d.CS + k) * d.nd;
        double* dst = d.sg + (int64_t(p) * d.MS + k) * d.nd;
        for (int c = blockIdx.x; c < d.nch; c += gridDim.x)
            for (int64_t e = int64_t(c) * d.chl + threadIdx.x, e1 = min(int64_t(c) * d.chl + d.chl, d.nd); e < e1;
                 e += BBS)
                dst[e] = src[e];
    }
}

__global__ void k_b_collapse_end(BDev d) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.P) return;
    BState& s = d.st[p];
    if (s.done || !s.collapse) return;
    if (s.added != s.cs) {  // "collapse_vectors generated less vectors than expected"
        s.status = -5;
        s.done = 1;
        s.conv = 0;
        return;
    }
    s.bs = s.cs;
    s.ss = s.cs;
    s.gv = 0;
    s.added = 0;
    s.collapse = 0;
    // the collapsed basis holds the Ritz vectors themselves: their coefficients are the identity
    for (int i = 0; i < BMS * BMS; ++i) s.alpha[i] = (i / BMS == i % BMS) ? 1.0 : 0.0;
}

// ---- Gram-Schmidt of one vector into the basis (add_row_and_orthonormalize) ---------------------
// kind 0: correction r_k (normalized), 1: collapse vector tmp_k, 2: random fill tmp_k, 3: guess
// tmp_k, 4: final Ritz vector tmp_k
__global__ void k_b_gs_begin(BDev d, int kind, int k) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.P) return;
    BState& s = d.st[p];
    s.gs_active = 0;
    s.gs_ok = 0;
    bool on;
    switch (kind) {
        case 0: on = !s.done && k < s.ntoadd; break;
        case 1: on = !s.done && s.collapse && k < s.cs; break;
        case 2: on = !s.done && k < s.nmiss; break;
        case 3: on = k < s.ntoadd; break;
        default: on = true; break;
    }
    if (!on) return;
    s.gs_active = 1;
    s.gs_scaled = 0;
    // rows to orthogonalize against: the basis so far (collapse / guess / final: rebuilt from 0)
    s.gs_row = (kind == 0 || kind == 2) ? s.bs + s.added : s.added;
    if (kind == 0) {
        s.gs_src = d.r + (int64_t(p) * d.NR + k) * d.nd;
        s.gs_a = s.rnorm[k] > 0.0 ? 1.0 / s.rnorm[k] : 0.0;  // normalize_vectors(r_)
    } else {
        s.gs_src = d.tmp + (int64_t(p) * d.TM + k) * d.nd;
        s.gs_a = 1.0;
    }
    for (int i = 0; i < BMS; ++i) s.gs_d[i] = 0.0;
}

// v = a src - sum_{i < row} d_i b_i into row `row`; partial |v|^2 and v . b_i, in one sweep: the
// basis elements of a position are loaded once into registers and serve both the subtraction and
// the overlaps of the result
__global__ __launch_bounds__(BBS) void k_b_gs_pass(BDev d) {
    __shared__ double red[BBS / 64][BMS + 1];
    __shared__ double dd[BMS];
    const int p = blockIdx.y;
    const BState& s = d.st[p];
    if (!s.gs_active) return;
    const int row = s.gs_row;
    for (int i = threadIdx.x; i < BMS; i += BBS) dd[i] = i < row ? s.gs_d[i] : 0.0;
    __syncthreads();
    const double a = s.gs_a;
    const double* src = s.gs_src;
    double* B = d.b + int64_t(p) * d.MS * d.nd;
    double* v = B + int64_t(row) * d.nd;
    double n2 = 0.0;
    double ov[BMS];
#pragma unroll
    for (int i = 0; i < BMS; ++i) ov[i] = 0.0;
    for (int c = blockIdx.x; c < d.nch; c += gridDim.x)
    for (int64_t e = int64_t(c) * d.chl + threadIdx.x, e1 = min(int64_t(c) * d.chl + d.chl, d.nd); e < e1; e += BBS) {
        double bv[BMS];
        double x = a * src[e];
#pragma unroll
        for (int i = 0; i < BMS; ++i) {
            bv[i] = i < row ? B[int64_t(i) * d.nd + e] : 0.0;
            x -= dd[i] * bv[i];
        }
        v[e] = x;
        n2 += x * x;
#pragma unroll
        for (int i = 0; i < BMS; ++i) ov[i] += x * bv[i];
    }
    // block sums of |v|^2 and the row overlaps
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    n2 = wave_sum(n2);
    if (lane == 0) red[w][0] = n2;
#pragma unroll
    for (int i = 0; i < BMS; ++i) {
        if (i >= row) break;
        const double t = wave_sum(ov[i]);
        if (lane == 0) red[w][1 + i] = t;
    }
    __syncthreads();
    double* out = d.part + (int64_t(p) * d.npb + blockIdx.x) * BPS;
    for (int i = threadIdx.x; i <= row; i += BBS) {
        double t = 0.0;
        for (int x = 0; x < BBS / 64; ++x) t += red[x][i];
        out[i] = t;
    }
}

// the reference's acceptance rules (discard 1e-7, orthogonality 1e-12), one wave per problem
__global__ __launch_bounds__(64) void k_b_gs_decide(BDev d) {
    __shared__ double sums[BMS + 1];
    const int p = blockIdx.x, l = threadIdx.x;
    BState& s = d.st[p];
    if (!s.gs_active) return;
    const int row = s.gs_row;
    if (l <= row) {
        double acc = 0.0;
        for (int c = 0; c < d.npb; ++c) acc += d.part[(int64_t(p) * d.npb + c) * BPS + l];
        sums[l] = acc;
    }
    __syncthreads();
    if (l != 0) return;
    const double norm2 = sums[0];
    double* B = d.b + int64_t(p) * d.MS * d.nd;
    if (s.gs_scaled) {  // this pass was the final scaling: its squared norm must be 1
        if (fabs(norm2 - 1.0) < d.orth) {
            s.gs_active = 0;
            s.gs_ok = 1;
            return;
        }
        s.gs_scaled = 0;
        s.gs_a = 1.0;
        for (int i = 0; i < row; ++i) s.gs_d[i] = 0.0;
        s.gs_src = B + int64_t(row) * d.nd;
        return;
    }
    const double normval = sqrt(norm2);
    if (!(normval >= d.discard)) {  // discarded (also for NaN)
        s.gs_active = 0;
        s.gs_ok = 0;
        return;
    }
    double mo = 0.0;
    for (int i = 0; i < row; ++i) mo = fmax(mo, fabs(sums[1 + i] / normval));
    s.gs_a = 1.0 / normval;
    s.gs_src = B + int64_t(row) * d.nd;
    if (mo < d.orth) {
        s.gs_scaled = 1;
        for (int i = 0; i < row; ++i) s.gs_d[i] = 0.0;
    } else {
        for (int i = 0; i < row; ++i) s.gs_d[i] = sums[1 + i] / normval;
    }
}

__global__ void k_b_gs_end(BDev d) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.P) return;
    BState& s = d.st[p];
    if (s.gs_ok) s.added += 1;
    s.gs_active = 0;  // not accepted within the passes: discarded
    s.gs_ok = 0;
}

__global__ void k_b_missing(BDev d) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.P) return;
    BState& s = d.st[p];
    if (s.done) return;
    s.nmiss = s.ntoadd - s.added;
}

// fill_random: tmp_k for k < nmiss from a SplitMix64 stream per problem
__global__ __launch_bounds__(BBS) void k_b_random(BDev d) {
    const int p = blockIdx.y;
    const BState& s = d.st[p];
    if (s.done || s.nmiss <= 0) return;
    for (int k = 0; k < min(s.nmiss, d.TM); ++k) {
        double* t = d.tmp + (int64_t(p) * d.TM + k) * d.nd;
        const unsigned long long seed = s.rng + 0x9E3779B97F4A7C15ULL * (unsigned long long)(s.iters * BNR + k + 1);
        for (int c = blockIdx.x; c < d.nch; c += gridDim.x)
        for (int64_t e = int64_t(c) * d.chl + threadIdx.x, e1 = min(int64_t(c) * d.chl + d.chl, d.nd); e < e1;
             e += BBS) {
            unsigned long long z = seed + 0x9E3779B97F4A7C15ULL * (unsigned long long)(e + 1);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
            z ^= z >> 31;
            t[e] = double(z >> 11) * (1.0 / 9007199254740992.0) - 0.5;
        }
    }
}

__global__ void k_b_iter_end(BDev d) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.P) return;
    BState& s = d.st[p];
    if (s.done) return;
    if (s.added == 0) {  // nothing could be added: end with the residual test's verdict
        s.done = 1;
        s.conv = s.r_ok;
        return;
    }
    s.bs += s.added;
    s.added = 0;
    for (int k = 0; k < BNR; ++k) s.lam_old[k] = s.lam[k];
}

__global__ void k_b_count_done(BDev d, int* out) {
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    int c = 0;
    for (int p = threadIdx.x; p < d.P; p += blockDim.x) c += d.st[p].done;
    atomicAdd(&cnt, c);
    __syncthreads();
    if (threadIdx.x == 0) *out = cnt;
}

// after the guess Gram-Schmidt: basis size = vectors accepted
__global__ void k_b_guess_end(BDev d, int should) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.P) return;
    BState& s = d.st[p];
    if (s.added != should) {  // "guess vectors are zero or linearly dependent"
        s.status = -5;
        s.done = 1;
        s.conv = 0;
    }
    s.bs = s.added;
    s.ss = 0;
    s.gv = 0;
    s.added = 0;
}

__global__ void k_b_final_begin(BDev d) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.P) return;
    d.st[p].added = 0;
}

static inline unsigned nb1(int64_t n, int bs) { return unsigned((n + bs - 1) / bs); }

// ---- one class of problems ------------------------------------------------------------------------
struct BatchClass {
    int n, na, nb, P, NR, MS, CS, NSL, ng;
    int64_t nd;
    SpinTables sa, sb;
    DBuf<uint32_t> binom;
    int gd[2] = {0, 0};
    DBuf<double> g2a, g2b, vpk, vpkh, hd, b, sg, r, tmp, tms, part, scratch;
    // Ms = 0 spin-flip symmetry (C^T = eta C): the Davidson-Liu vectors are packed coordinates
    // (nd of them) of full vectors of ndf = N^2 determinants, and sigma has no beta same-spin term
    int eta = 0;
    int64_t N = 0, ndf = 0;
    DBuf<BState> st;
    DBuf<int> pvo, done_dev;
    DBuf<int64_t> pvoff, pidx;
    DBuf<double> pval;
    // sigma tables
    DBuf<const double*> c_tab, g2a_tab, g2b_tab, vpk_tab, vpkh_tab;
    DBuf<double*> s_tab, ct_tab, yt_tab, e_tab;
    hipStream_t stream = nullptr;
    int* done_host = nullptr;
    struct Variant {
        int64_t chl;
        int sb;
        hipGraphExec_t exec = nullptr;
        double ms = 0.0;
    };
    std::vector<Variant> var;
    int chosen = 0;
    bool graphs = true;
    void launch(int vi) {
        if (graphs)
            HIPCHK(hipGraphLaunch(var[vi].exec, stream));
        else
            iteration(var[vi].chl, var[vi].sb);
    }
    std::vector<double> ecore;
    int iters_run = 0;
    bool active = true;

    BDev dev(int64_t chl) {
        BDev d;
        d.P = P; d.NR = NR; d.MS = MS; d.CS = CS; d.NSL = NSL;
        d.nd = nd;
        d.chl = chl;
        d.nch = int((nd + chl - 1) / chl);
        d.npb = std::min(d.nch, BNPB);
        d.st = st.p;
        d.b = b.p; d.sg = sg.p; d.r = r.p; d.tmp = tmp.p; d.tms = tms.p;
        d.TM = std::max(CS, std::max(NR, ng));
        d.hd = hd.p;
        d.part = part.p;
        d.pvo = pvo.p; d.pvoff = pvoff.p; d.pidx = pidx.p; d.pval = pval.p;
        d.e_conv = e_conv; d.r_conv = r_conv; d.discard = 1.0e-7; d.orth = 1.0e-12;
        d.maxiter = maxiter;
        return d;
    }
    double e_conv = 1e-12, r_conv = 1e-6;
    int maxiter = 100;
    int64_t min_chl = 0;

    ~BatchClass() {
        for (auto& v : var)
            if (v.exec) (void)hipGraphExecDestroy(v.exec);
        if (done_host) (void)hipHostFree(done_host);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // Gram-Schmidt sequence of one vector (BGS_PASSES passes) for all problems
    void gs_vector(const BDev& d, int kind, int k) {
        const unsigned gp = nb1(P, 64);
        const int passes = kind == 1 ? BGS_PASSES_COLL : (kind == 0 || kind == 2) ? BGS_PASSES_CORR : BGS_PASSES;
        k_b_gs_begin<<<gp, 64, 0, stream>>>(d, kind, k);
        for (int it = 0; it < passes; ++it) {
            k_b_gs_pass<<<dim3(d.npb, P), BBS, 0, stream>>>(d);
            k_b_gs_decide<<<P, 64, 0, stream>>>(d);
        }
        k_b_gs_end<<<gp, 64, 0, stream>>>(d);
    }

    void sigma_items(int sbsz) {
        const int nitem = P * NSL;
        for (int i0 = 0; i0 < nitem; i0 += sbsz) {
            const int m = std::min(sbsz, nitem - i0);
            BatchSigmaTabs t;
            t.nitem = m;
            t.act = reinterpret_cast<const void* const*>(c_tab.p + i0);
            t.c = c_tab.p + i0;
            t.s = s_tab.p + i0;
            t.ct = ct_tab.p;  // scratch slots 0..m-1, reused by every sub-batch
            t.yt = yt_tab.p;
            t.e = e_tab.p;
            t.g2a = g2a_tab.p + i0;
            t.g2b = g2b_tab.p + i0;
            if (eta) {
                t.vpk = vpkh_tab.p + i0;
                batch_sigma_f64_ms0(sa, sb, gd[0], eta, t, stream);
            } else {
                t.vpk = vpk_tab.p + i0;
                batch_sigma_f64(sa, sb, gd[0], gd[1], t, stream);
            }
        }
    }

    // one Davidson-Liu iteration of every problem of the class
    void iteration(int64_t chl, int sbsz) {
        const BDev d = dev(chl);
        const unsigned gp = nb1(P, 64);
        const dim3 gs(d.npb, P);
        k_b_sig_tab<<<nb1(int64_t(P) * NSL, 64), 64, 0, stream>>>(d, c_tab.p, s_tab.p);
        sigma_items(sbsz);
        k_b_after_sigma<<<gp, 64, 0, stream>>>(d);
        k_b_gdots<<<gs, BBS, 0, stream>>>(d);
        k_b_eigh<<<P, 64, 0, stream>>>(d);
        k_b_residual<<<gs, BBS, 0, stream>>>(d);
        k_b_converge<<<gp, 64, 0, stream>>>(d);
        k_b_project<<<dim3(P, NR), BBS, 0, stream>>>(d, d.r, NR, 0);
        k_b_rnorm_part<<<gs, BBS, 0, stream>>>(d);
        k_b_rnorm<<<gp, 64, 0, stream>>>(d);
        // collapse: b <- orthonormalized alpha^T b, sigma <- alpha^T sigma
        k_b_rotate<<<gs, BBS, 0, stream>>>(d, d.b, MS, d.tmp, d.TM, 0);
        k_b_collapse_begin<<<gp, 64, 0, stream>>>(d);
        for (int k = 0; k < CS; ++k) gs_vector(d, 1, k);
        k_b_rotate<<<gs, BBS, 0, stream>>>(d, d.sg, MS, d.tms, CS, 0);
        k_b_copy_sigma<<<gs, BBS, 0, stream>>>(d);
        k_b_collapse_end<<<gp, 64, 0, stream>>>(d);
        // correction vectors, then random vectors for any that were discarded
        for (int k = 0; k < NR; ++k) gs_vector(d, 0, k);
        k_b_missing<<<gp, 64, 0, stream>>>(d);
        k_b_random<<<gs, BBS, 0, stream>>>(d);
        k_b_project<<<dim3(P, NR), BBS, 0, stream>>>(d, d.tmp, d.TM, 1);
        for (int k = 0; k < NR; ++k) gs_vector(d, 2, k);
        k_b_iter_end<<<gp, 64, 0, stream>>>(d);
        k_b_count_done<<<1, 256, 0, stream>>>(d, done_dev.p);
        HIPCHK(hipGetLastError());
    }
};

}  // namespace fci

// ===================================================================================================
// C ABI
// ===================================================================================================
using fci::Error;

namespace {
thread_local std::string g_batch_err;
}

extern "C" {

const char* fci_batch_last_error(void) { return g_batch_err.c_str(); }

int fci_batch_solve(int nprob, const int* norb, const int* nalpha, const int* nbeta,
                    const double* const* oei_a, const double* const* oei_b, const double* const* tei_aa,
                    const double* const* tei_ab, const double* const* tei_bb, const double* e_core,
                    const fci_options* opts, double* energies, int* iterations, int* converged,
                    double* const* evecs, fci_batch_stats* stats) {
    using namespace fci;
    if (nprob < 0 || (nprob > 0 && (!norb || !nalpha || !nbeta || !oei_a || !oei_b || !tei_aa || !tei_ab || !tei_bb ||
                                    !opts || !energies)))
        return FCI_ERR_ARG;
    try {
        check_device_public();
        const auto t_start = std::chrono::steady_clock::now();
        const int NR = opts->nroot;
        if (NR < 1 || NR > BNR) throw Error(FCI_ERR_ARG, "batched solves support 1 to 4 roots");
        const int MS0 = opts->subspace_per_root * NR;
        if (MS0 > BMS) throw Error(FCI_ERR_ARG, "batched solves support subspaces of at most 32 vectors");
        if (opts->precision != FCI_PREC_FP64 || opts->preconditioner != FCI_PRECOND_DIAGONAL)
            throw Error(FCI_ERR_ARG, "batched solves run in FP64 with the diagonal preconditioner");
        if (opts->spin_adapt) throw Error(FCI_ERR_ARG, "batched solves run in the determinant basis (no CI_SPIN_ADAPT)");
        // group problems by shape
        std::map<std::tuple<int, int, int>, std::vector<int>> groups;
        for (int i = 0; i < nprob; ++i) {
            if (norb[i] < 2 || norb[i] > kMaxOrb || nalpha[i] < 0 || nbeta[i] < 0 || nalpha[i] > norb[i] ||
                nbeta[i] > norb[i])
                throw Error(FCI_ERR_ARG, "invalid problem shape");
            groups[{norb[i], nalpha[i], nbeta[i]}].push_back(i);
        }
        std::vector<std::unique_ptr<BatchClass>> classes;
        double setup_s = 0.0;
        for (auto& [shape, ids] : groups) {
            auto c = std::make_unique<BatchClass>();
            BatchClass& C = *c;
            C.n = std::get<0>(shape);
            C.na = std::get<1>(shape);
            C.nb = std::get<2>(shape);
            C.P = int(ids.size());
            C.NR = NR;
            auto bt = binom_table();
            C.binom.alloc(bt.size());
            C.binom.upload(bt.data(), bt.size());
            HIPCHK(hipStreamCreateWithFlags(&C.stream, hipStreamNonBlocking));
            build_spin_tables(C.sa, C.n, C.na, C.binom.p, C.stream);
            build_spin_tables(C.sb, C.n, C.nb, C.binom.p, C.stream);
            HIPCHK(hipStreamSynchronize(C.stream));
            std::string why;
            if (!batch_sigma_supported(C.sa, C.sb, &why)) throw Error(FCI_ERR_ARG, why);
            C.ndf = C.sa.nstr * C.sb.nstr;
            C.N = C.sa.nstr;
            {
                const size_t m2 = size_t(C.n) * C.n, m4 = m2 * m2;
                bool same = opts->ms0_symmetry && C.na == C.nb && opts->multiplicity % 2 == 1 && C.ndf >= 2;
                for (int i : ids) {
                    if (!same) break;
                    same = std::memcmp(oei_a[i], oei_b[i], m2 * sizeof(double)) == 0 &&
                           std::memcmp(tei_aa[i], tei_bb[i], m4 * sizeof(double)) == 0;
                }
                C.eta = same ? (((opts->multiplicity - 1) / 2) % 2 == 0 ? 1 : -1) : 0;
                if (C.eta && ms0_packed_size(C.N, C.eta) < 4 * int64_t(NR)) C.eta = 0;  // tiny spaces
            }
            C.nd = C.eta ? ms0_packed_size(C.N, C.eta) : C.ndf;
            C.MS = int(std::min<int64_t>(MS0, C.nd));
            C.CS = int(std::min<int64_t>(std::max(opts->collapse_per_root * NR, NR), C.MS));
            C.ng = int(std::min<int64_t>(int64_t(opts->guess_per_root) * NR, C.nd));
            if (C.ng < NR || C.ng > C.MS) throw Error(FCI_ERR_ARG, "invalid number of guess vectors");
            C.NSL = std::max(NR, C.ng);
            C.e_conv = opts->e_convergence;
            C.r_conv = opts->r_convergence;
            C.maxiter = opts->maxiter;
            const int P = C.P;
            const int64_t nd = C.nd;
            const int TM = std::max(C.CS, std::max(NR, C.ng));
            // integrals, H diagonal, guesses
            const int n = C.n;
            const size_t n2 = size_t(n) * n, n4 = n2 * n2;
            std::vector<double> g2a_h, g2b_h, vpk_h;
            C.hd.alloc(size_t(P) * nd);
            DBuf<double> hfull;  // the full H diagonal of one problem (symmetric classes)
            if (C.eta) hfull.alloc(C.ndf);
            DBuf<double> oa, ob, taa, tab, tbb;
            oa.alloc(n2); ob.alloc(n2); taa.alloc(n4); tab.alloc(n4); tbb.alloc(n4);
            std::vector<std::vector<SparseVec>> guesses(P), bad(P);
            C.ecore.resize(P);
            auto host_ints = [&](int i) {
                HostInts hi;
                hi.n = n;
                hi.oei_a.assign(oei_a[i], oei_a[i] + n2);
                hi.oei_b.assign(oei_b[i], oei_b[i] + n2);
                hi.tei_aa.assign(tei_aa[i], tei_aa[i] + n4);
                hi.tei_ab.assign(tei_ab[i], tei_ab[i] + n4);
                hi.tei_bb.assign(tei_bb[i], tei_bb[i] + n4);
                hi.e_core = 0.0;  // solved without the constant; added to the energies at the end
                return hi;
            };
            // device part, problem by problem on the class stream: packed integrals, H diagonal,
            // the determinants of the guess space (one set of selection buffers for all)
            const int ngd = int(std::min<int64_t>(int64_t(C.ng) * opts->ndets_per_guess, C.ndf));
            std::vector<std::vector<int64_t>> sel_idx(P);
            SelectScratch sel;
            for (int q = 0; q < P; ++q) {
                const int i = ids[q];
                C.ecore[q] = e_core ? e_core[i] : 0.0;
                PackedInts pk = pack_integrals(host_ints(i), C.sa, C.sb);
                if (q == 0) {
                    C.gd[0] = pk.gd[0];
                    C.gd[1] = pk.gd[1];
                }
                g2a_h.insert(g2a_h.end(), pk.g2[0].begin(), pk.g2[0].end());
                g2b_h.insert(g2b_h.end(), pk.g2[1].begin(), pk.g2[1].end());
                vpk_h.insert(vpk_h.end(), pk.vpk.begin(), pk.vpk.end());
                oa.upload(oei_a[i], n2); ob.upload(oei_b[i], n2);
                taa.upload(tei_aa[i], n4); tab.upload(tei_ab[i], n4); tbb.upload(tei_bb[i], n4);
                double* hdq = C.eta ? hfull.p : C.hd.p + size_t(q) * nd;
                hdiag_device(C.sa, C.sb, oa.p, ob.p, taa.p, tab.p, tbb.p, 0.0, hdq, C.stream);
                std::vector<double> vals;
                smallest_k(hdq, C.ndf, ngd, vals, sel_idx[q], C.stream, &sel);  // ends synchronized
                // the diagonal in the packed basis (a determinant and its mirror averaged); the
                // next problem's diagonal is written after this on the same stream
                if (C.eta)
                    ms0_pack_device<double, double>(hfull.p, C.hd.p + size_t(q) * nd, C.N, C.eta, 1, 0.5, 1.0,
                                                    C.stream);
            }
            HIPCHK(hipStreamSynchronize(C.stream));
            // host part, problems over threads: the guess vectors and the roots of other
            // multiplicities to project out (in packed coordinates for symmetric classes, where
            // roots of the other spin parity vanish)
            {
                std::atomic<int> next{0};
                std::vector<std::exception_ptr> err(P);
                auto work = [&] {
                    for (int q; (q = next.fetch_add(1)) < P;) {
                        try {
                            auto gb = initial_guess_from_selection(sel_idx[q], C.sb.nstr, n, C.na, C.nb,
                                                                   host_ints(ids[q]), opts->multiplicity, C.ng);
                            guesses[q] = std::move(gb.first);
                            bad[q] = std::move(gb.second);
                            if (!C.eta) continue;
                            for (auto& g : guesses[q]) g = ms0_pack_sparse(g, C.N, C.eta);
                            std::vector<SparseVec> keep;
                            for (const auto& v : bad[q]) {
                                SparseVec w = ms0_pack_sparse(v, C.N, C.eta);
                                double w2 = 0.0;
                                for (const auto& e : w) w2 += e.second * e.second;
                                if (w2 < 1e-12) continue;
                                for (auto& e : w) e.second /= std::sqrt(w2);
                                keep.push_back(std::move(w));
                            }
                            bad[q] = std::move(keep);
                        } catch (...) {
                            err[q] = std::current_exception();
                        }
                    }
                };
                const int nth = std::max(1, std::min<int>({P, 16, int(std::thread::hardware_concurrency())}));
                std::vector<std::thread> pool;
                try {
                    for (int t = 1; t < nth; ++t) pool.emplace_back(work);
                } catch (const std::system_error&) {
                    // fewer threads than asked for: the ones running and this one share the rest
                }
                work();
                for (auto& t : pool) t.join();
                for (auto& e : err)
                    if (e) std::rethrow_exception(e);
            }
            C.g2a.alloc(g2a_h.size()); C.g2a.upload(g2a_h.data(), g2a_h.size());
            C.g2b.alloc(g2b_h.size()); C.g2b.upload(g2b_h.data(), g2b_h.size());
            C.vpk.alloc(vpk_h.size()); C.vpk.upload(vpk_h.data(), vpk_h.size());
            if (C.eta) {  // halved packed integrals (exact) for the symmetric sigma
                for (double& x : vpk_h) x *= 0.5;
                C.vpkh.alloc(vpk_h.size());
                C.vpkh.upload(vpk_h.data(), vpk_h.size());
            }
            // vectors and state
            C.b.alloc(size_t(P) * C.MS * nd);
            C.sg.alloc(size_t(P) * C.MS * nd);
            C.r.alloc(size_t(P) * NR * nd);
            C.tmp.alloc(size_t(P) * TM * nd);
            C.tms.alloc(size_t(P) * C.CS * nd);
            C.st.alloc(P);
            C.done_dev.alloc(1);
            HIPCHK(hipHostMalloc((void**)&C.done_host, sizeof(int)));
            // autotuner candidates: streaming chunk length x sigma sub-batch size
            std::vector<int64_t> chls;
            for (int64_t x : {int64_t(4096), int64_t(16384), int64_t(65536)})
                if (x < nd || chls.empty()) chls.push_back(std::min<int64_t>(x, std::max<int64_t>(nd, 1)));
            chls.erase(std::unique(chls.begin(), chls.end()), chls.end());
            const int nitem = P * C.NSL;
            std::vector<int> sbs = {nitem};
            size_t oyt = 0, oe = 0;
            const size_t scr = batch_sigma_scratch_elems(C.sa, C.sb, &oyt, &oe);
            if (nitem > 16 && scr * 8 * size_t(nitem) > (size_t(8) << 30)) sbs.push_back(16);
            for (int64_t x : chls)
                for (int y : sbs) C.var.push_back({x, y});
            C.min_chl = chls.front();
            C.part.alloc(size_t(P) * BNPB * BPS);
            const int slots = *std::max_element(sbs.begin(), sbs.end());
            C.scratch.alloc(scr * size_t(slots));
            // bad roots
            std::vector<int> pvo(P + 1, 0);
            std::vector<int64_t> pvoff(1, 0), pidx;
            std::vector<double> pval;
            for (int q = 0; q < P; ++q) {
                for (const auto& v : bad[q]) {
                    for (const auto& [ix, val] : v) {
                        pidx.push_back(ix);
                        pval.push_back(val);
                    }
                    pvoff.push_back(int64_t(pidx.size()));
                }
                pvo[q + 1] = int(pvoff.size()) - 1;
            }
            if (pidx.empty()) {
                pidx.push_back(0);
                pval.push_back(0.0);
            }
            C.pvo.alloc(pvo.size()); C.pvo.upload(pvo.data(), pvo.size());
            C.pvoff.alloc(pvoff.size()); C.pvoff.upload(pvoff.data(), pvoff.size());
            C.pidx.alloc(pidx.size()); C.pidx.upload(pidx.data(), pidx.size());
            C.pval.alloc(pval.size()); C.pval.upload(pval.data(), pval.size());
            // sigma tables: per item integrals, per slot scratch
            std::vector<const double*> ga(nitem), gb2(nitem), vp(nitem), vph(nitem);
            for (int it = 0; it < nitem; ++it) {
                const int q = it / C.NSL;
                ga[it] = C.g2a.p + size_t(q) * C.gd[0] * C.gd[0];
                gb2[it] = C.g2b.p + size_t(q) * C.gd[1] * C.gd[1];
                vp[it] = C.vpk.p + size_t(q) * size_t(n * (n + 1) / 2) * size_t(n * (n + 1) / 2);
                vph[it] = C.eta ? C.vpkh.p + size_t(q) * size_t(n * (n + 1) / 2) * size_t(n * (n + 1) / 2) : nullptr;
            }
            std::vector<double*> ct(slots), yt(slots), et(slots);
            for (int s = 0; s < slots; ++s) {
                ct[s] = C.scratch.p + size_t(s) * scr;
                yt[s] = ct[s] + oyt;
                et[s] = ct[s] + oe;
            }
            C.c_tab.alloc(nitem); C.s_tab.alloc(nitem);
            C.g2a_tab.alloc(nitem); C.g2a_tab.upload(ga.data(), nitem);
            C.g2b_tab.alloc(nitem); C.g2b_tab.upload(gb2.data(), nitem);
            C.vpk_tab.alloc(nitem); C.vpk_tab.upload(vp.data(), nitem);
            C.vpkh_tab.alloc(nitem); C.vpkh_tab.upload(vph.data(), nitem);
            C.ct_tab.alloc(slots); C.ct_tab.upload(ct.data(), slots);
            C.yt_tab.alloc(slots); C.yt_tab.upload(yt.data(), slots);
            C.e_tab.alloc(slots); C.e_tab.upload(et.data(), slots);
            // initial basis: guesses -> tmp (one copy of all entries), project, Gram-Schmidt
            std::vector<int64_t> goff(1, 0), gidx;
            std::vector<double> gval;
            for (int q = 0; q < P; ++q) {
                if (int(guesses[q].size()) != C.ng)
                    throw Error(FCI_ERR_ARG, "guess construction returned an unexpected number of vectors");
                for (const auto& g : guesses[q]) {
                    for (const auto& [ix, val] : g) {
                        gidx.push_back(ix);
                        gval.push_back(val);
                    }
                    goff.push_back(int64_t(gidx.size()));
                }
            }
            if (gidx.empty()) {
                gidx.push_back(0);
                gval.push_back(0.0);
            }
            DBuf<int64_t> dgoff, dgidx;
            DBuf<double> dgval;
            dgoff.alloc(goff.size()); dgoff.upload(goff.data(), goff.size());
            dgidx.alloc(gidx.size()); dgidx.upload(gidx.data(), gidx.size());
            dgval.alloc(gval.size()); dgval.upload(gval.data(), gval.size());
            const BDev d = C.dev(C.min_chl);
            k_b_zero<<<4096, 256, 0, C.stream>>>(C.tmp.p, int64_t(P) * TM * nd);
            k_b_init<<<nb1(P, 64), 64, 0, C.stream>>>(d, C.ng);
            k_b_scatter_guess<<<dim3(1, P, C.ng), 256, 0, C.stream>>>(d, C.ng, dgoff.p, dgidx.p, dgval.p);
            k_b_project<<<dim3(P, std::min(NR, C.ng)), BBS, 0, C.stream>>>(d, d.tmp, d.TM, 0);
            k_b_final_begin<<<nb1(P, 64), 64, 0, C.stream>>>(d);
            for (int k = 0; k < C.ng; ++k) C.gs_vector(d, 3, k);
            k_b_guess_end<<<nb1(P, 64), 64, 0, C.stream>>>(d, C.ng);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(C.stream));
            // capture one iteration per candidate (kernel attributes are set outside the capture);
            // FCI_BATCH_GRAPHS=0 launches the same sequence directly each iteration instead
            // (profilers that cannot follow graph replays)
            batch_sigma_prepare(C.sa, C.sb);
            static const bool use_graphs = [] {
                const char* e = std::getenv("FCI_BATCH_GRAPHS");
                return !e || std::atoi(e) != 0;
            }();
            C.graphs = use_graphs;
            if (use_graphs)
            for (auto& v : C.var) {
                hipGraph_t g;
                HIPCHK(hipStreamBeginCapture(C.stream, hipStreamCaptureModeThreadLocal));
                C.iteration(v.chl, v.sb);
                HIPCHK(hipStreamEndCapture(C.stream, &g));
                HIPCHK(hipGraphInstantiate(&v.exec, g, nullptr, nullptr, 0));
                HIPCHK(hipGraphDestroy(g));
            }
            classes.push_back(std::move(c));
        }
        setup_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
        // iterate: the classes' streams run side by side; the host polls each class's done count
        const auto t_iter = std::chrono::steady_clock::now();
        constexpr int kPoll = 4;
        int max_it = 0;
        for (auto& c : classes) max_it = std::max(max_it, c->maxiter);
        // autotune: after one untimed iteration (the guess vectors' sigma builds make the first one
        // heavier than the rest), the candidates run one iteration each in turn, twice over, timed
        // with events; a candidate's time is its faster round (iterations differ in basis size
        // and collapses, so interleaving keeps the comparison fair)
        for (auto& c : classes) {
            BatchClass& C = *c;
            if (C.var.size() < 2) continue;
            hipEvent_t e0, e1;
            HIPCHK(hipEventCreate(&e0));
            HIPCHK(hipEventCreate(&e1));
            C.launch(0);
            C.iters_run += 1;
            for (auto& v : C.var) v.ms = 1e30;
            for (int round = 0; round < 2; ++round)
                for (auto& v : C.var) {
                    HIPCHK(hipEventRecord(e0, C.stream));
                    C.launch(int(&v - C.var.data()));
                    HIPCHK(hipEventRecord(e1, C.stream));
                    HIPCHK(hipEventSynchronize(e1));
                    float ms = 0.f;
                    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                    v.ms = std::min(v.ms, double(ms));
                    C.iters_run += 1;
                }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            int best = 0;
            for (int i = 1; i < int(C.var.size()); ++i)
                if (C.var[i].ms < C.var[best].ms) best = i;
            C.chosen = best;
        }
        for (;;) {
            bool any = false;
            for (auto& c : classes) {
                BatchClass& C = *c;
                if (!C.active) continue;
                any = true;
                for (int r = 0; r < kPoll && C.iters_run < C.maxiter; ++r, ++C.iters_run)
                    C.launch(C.chosen);
                HIPCHK(hipMemcpyAsync(C.done_host, C.done_dev.p, sizeof(int), hipMemcpyDeviceToHost, C.stream));
            }
            if (!any) break;
            for (auto& c : classes) {
                BatchClass& C = *c;
                if (!C.active) continue;
                HIPCHK(hipStreamSynchronize(C.stream));
                if (*C.done_host >= C.P || C.iters_run >= C.maxiter) C.active = false;
            }
        }
        (void)max_it;
        const double iter_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_iter).count();
        // results: final Ritz vectors (get_results), energies, counts
        int nconv = 0, total_iters = 0, total_sigma = 0;
        for (auto& c : classes) {
            BatchClass& C = *c;
            const BDev d = C.dev(C.min_chl);
            const dim3 gs(d.npb, C.P);
            // get_results (davidson_liu_solver.cc:488-501): Ritz vectors, orthonormalized into b
            k_b_rotate<<<gs, BBS, 0, C.stream>>>(d, d.b, C.MS, d.tmp, d.TM, 1);
            k_b_final_begin<<<nb1(C.P, 64), 64, 0, C.stream>>>(d);
            for (int k = 0; k < NR; ++k) C.gs_vector(d, 4, k);
            HIPCHK(hipGetLastError());
            std::vector<BState> hs(C.P);
            HIPCHK(hipMemcpyAsync(hs.data(), C.st.p, sizeof(BState) * C.P, hipMemcpyDeviceToHost, C.stream));
            HIPCHK(hipStreamSynchronize(C.stream));
            const auto& ids = groups[{C.n, C.na, C.nb}];
            for (int q = 0; q < C.P; ++q) {
                const int i = ids[q];
                const BState& s = hs[q];
                if (s.status < 0) throw Error(s.status, "Davidson-Liu numerical failure in a batched problem");
                for (int k = 0; k < NR; ++k) energies[size_t(i) * NR + k] = s.lam[k] + C.ecore[q];
                if (evecs && evecs[i]) {
                    const double* bq = C.b.p + size_t(q) * C.MS * C.nd;
                    if (C.eta) {  // determinant vectors of the packed ones
                        DBuf<double> full(C.ndf);
                        for (int k = 0; k < NR; ++k) {
                            ms0_unpack_device<double, double>(bq + size_t(k) * C.nd, full.p, C.N, C.eta, C.stream);
                            HIPCHK(hipStreamSynchronize(C.stream));
                            HIPCHK(hipMemcpy(evecs[i] + size_t(k) * C.ndf, full.p, sizeof(double) * C.ndf,
                                             hipMemcpyDeviceToHost));
                        }
                    } else {
                        HIPCHK(hipMemcpy(evecs[i], bq, sizeof(double) * NR * C.nd, hipMemcpyDeviceToHost));
                    }
                }
                if (iterations) iterations[i] = s.iters;
                if (converged) converged[i] = s.conv;
                nconv += s.conv;
                total_iters += s.iters;
                total_sigma += s.nsigma;
            }
        }
        if (stats) {
            stats->setup_seconds = setup_s;
            stats->iterate_seconds = iter_s;
            stats->nclasses = int(classes.size());
            stats->converged = nconv;
            stats->iterations = total_iters;
            stats->sigma_builds = total_sigma;
            int k = 0;
            for (auto& c : classes) {
                if (k >= FCI_BATCH_MAX_CLASSES) break;
                stats->class_norb[k] = c->n;
                stats->class_nprob[k] = c->P;
                stats->class_chunk[k] = int(c->var[c->chosen].chl);
                stats->class_subbatch[k] = c->var[c->chosen].sb;
                stats->class_graph_launches[k] = c->iters_run;
                stats->class_ms_per_iteration[k] = c->var[c->chosen].ms;
                ++k;
            }
        }
        return FCI_OK;
    } catch (const Error& e) {
        g_batch_err = e.what();
        return e.status;
    } catch (const std::exception& e) {
        g_batch_err = e.what();
        return FCI_ERR_ARG;
    }
}

}  // extern "C"
