// fci_dl.hip -- Davidson-Liu on the device, following forte/helpers/davidson_liu_solver.cc.
// The subspace problem (at most a few dozen vectors) is diagonalized on the host; every
// operation on vectors of the full space runs in FP64 on the GPU.
#include "fci_dl.h"

#include <type_traits>

#include <algorithm>
#include <cmath>
#include <cstring>

namespace fci {

static inline unsigned nblk(int64_t n, int bs) { return unsigned((n + bs - 1) / bs); }

void jacobi_eigh(int n, const std::vector<double>& A0, std::vector<double>& evals, std::vector<double>& Vout) {
    std::vector<double> A(A0), V(size_t(n) * n, 0.0);
    for (int i = 0; i < n; ++i) V[size_t(i) * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                double a = A[size_t(i) * n + j];
                tot += a * a;
                if (i != j) off += a * a;
            }
        if (off == 0.0 || off <= 1e-30 * (tot > 0 ? tot : 1.0)) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = A[size_t(p) * n + q];
                if (std::fabs(apq) < 1e-300) continue;
                double app = A[size_t(p) * n + p], aqq = A[size_t(q) * n + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    double akp = A[size_t(k) * n + p], akq = A[size_t(k) * n + q];
                    A[size_t(k) * n + p] = c * akp - s * akq;
                    A[size_t(k) * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    double apk = A[size_t(p) * n + k], aqk = A[size_t(q) * n + k];
                    A[size_t(p) * n + k] = c * apk - s * aqk;
                    A[size_t(q) * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    double vkp = V[size_t(k) * n + p], vkq = V[size_t(k) * n + q];
                    V[size_t(k) * n + p] = c * vkp - s * vkq;
                    V[size_t(k) * n + q] = s * vkp + c * vkq;
                }
            }
    }
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return A[size_t(a) * n + a] < A[size_t(b) * n + b]; });
    evals.resize(n);
    Vout.assign(size_t(n) * n, 0.0);
    for (int k = 0; k < n; ++k) {
        evals[k] = A[size_t(idx[k]) * n + idx[k]];
        for (int i = 0; i < n; ++i) Vout[size_t(k) * n + i] = V[size_t(i) * n + idx[k]];
    }
}

// ---- kernels -----------------------------------------------------------------------------------
constexpr int DL_BS = 256;
constexpr int DL_MAXM = 64;
constexpr int DL_NBLK = 1024;  // blocks of the reduction kernels

// out_part[blk][i] = sum over this block's elements of x[e] * Y_i[e]  (i < m)
template <class TX, class TY>  // FP64 or FP32 vectors (FP32 sigma rows of the mixed phase), FP64 sums
__global__ __launch_bounds__(DL_BS) void k_dots(const TX* __restrict__ x, const TY* __restrict__ Y,
                                                int m, int64_t n, double* __restrict__ part) {
    __shared__ double red[DL_MAXM][DL_BS / 64];
    double acc[DL_MAXM];
#pragma unroll
    for (int i = 0; i < DL_MAXM; ++i) acc[i] = 0.0;
    for (int64_t e = int64_t(blockIdx.x) * DL_BS + threadIdx.x; e < n; e += int64_t(gridDim.x) * DL_BS) {
        double xv = double(x[e]);
#pragma unroll
        for (int i = 0; i < DL_MAXM; ++i)
            if (i < m) acc[i] += xv * double(Y[size_t(i) * n + e]);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < DL_MAXM; ++i) {
        if (i >= m) break;
        double v = acc[i];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) red[i][wave] = v;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < m; i += DL_BS) {
        double s = 0.0;
        for (int w = 0; w < DL_BS / 64; ++w) s += red[i][w];
        part[size_t(blockIdx.x) * m + i] = s;
    }
}

// part[blk][i] for the vectors i of chunk blockIdx.y (8 per chunk), double2 loads (x and every
// Y_i 16-byte aligned: n even)
constexpr int DOT_MC = 8;
template <class T> struct Vec2;
template <> struct Vec2<double> { using type = double2; };
template <> struct Vec2<float> { using type = float2; };
template <class TX, class TY>
__global__ __launch_bounds__(DL_BS) void k_dots8(const TX* __restrict__ x, const TY* __restrict__ Y, int m,
                                                 int64_t n, double* __restrict__ part) {
    using X2 = typename Vec2<TX>::type;
    using Y2 = typename Vec2<TY>::type;
    __shared__ double red[DOT_MC][DL_BS / 64];
    const int i0 = blockIdx.y * DOT_MC;
    const int mc = min(DOT_MC, m - i0);
    double acc[DOT_MC];
#pragma unroll
    for (int i = 0; i < DOT_MC; ++i) acc[i] = 0.0;
    const int64_t n2 = n / 2;
    const X2* x2 = reinterpret_cast<const X2*>(x);
    for (int64_t e = int64_t(blockIdx.x) * DL_BS + threadIdx.x; e < n2; e += int64_t(gridDim.x) * DL_BS) {
        const X2 xv = x2[e];
#pragma unroll
        for (int i = 0; i < DOT_MC; ++i)
            if (i < mc) {
                const Y2 y = reinterpret_cast<const Y2*>(Y + size_t(i0 + i) * n)[e];
                acc[i] += double(xv.x) * double(y.x) + double(xv.y) * double(y.y);
            }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < DOT_MC; ++i) {
        double v = acc[i];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) red[i][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < mc) {
        double s = 0.0;
        for (int w = 0; w < DL_BS / 64; ++w) s += red[threadIdx.x][w];
        part[size_t(blockIdx.x) * m + i0 + threadIdx.x] = s;
    }
}

// Fused classical Gram-Schmidt step: v <- a src - sum_i d_i A_i (i < m; src may be v), then the
// partial sums of ||v||^2 (slot m) and v . A_i (slots i < m) of the updated v, in one pass over A.
constexpr int GS_MAXM = 32;
template <int GS_M>  // compile-time bound on m: the accumulators live in registers
__global__ __launch_bounds__(DL_BS) void k_gs_fused(double* v, const double* src, const double* __restrict__ A,
                                                    int m, const double* __restrict__ d, double a, int64_t n,
                                                    double* __restrict__ part) {
    __shared__ double dd[GS_M];
    __shared__ double red[GS_M + 1][DL_BS / 64];
    for (int i = threadIdx.x; i < m; i += DL_BS) dd[i] = d[i];
    __syncthreads();
    double acc[GS_M + 1];
#pragma unroll
    for (int i = 0; i <= GS_M; ++i) acc[i] = 0.0;
    for (int64_t e = int64_t(blockIdx.x) * DL_BS + threadIdx.x; e < n; e += int64_t(gridDim.x) * DL_BS) {
        double av[GS_M];
        double x = a * src[e];
#pragma unroll
        for (int i = 0; i < GS_M; ++i)
            if (i < m) {
                av[i] = A[size_t(i) * n + e];
                x -= dd[i] * av[i];
            }
        v[e] = x;
        acc[GS_M] += x * x;
#pragma unroll
        for (int i = 0; i < GS_M; ++i)
            if (i < m) acc[i] += x * av[i];
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i <= GS_M; ++i) {
        if (i < m || i == GS_M) {
            double t = acc[i];
            for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
            if (lane == 0) red[i][wave] = t;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i <= m; i += DL_BS) {
        const int src = i < m ? i : GS_M;
        double t = 0.0;
        for (int w = 0; w < DL_BS / 64; ++w) t += red[src][w];
        part[size_t(blockIdx.x) * (m + 1) + i] = t;
    }
}

__global__ void k_reduce_parts(const double* __restrict__ part, int nparts, int m, double* __restrict__ out) {
    int i = blockIdx.x;
    __shared__ double red[256];
    double s = 0.0;
    for (int p = threadIdx.x; p < nparts; p += blockDim.x) s += part[size_t(p) * m + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[i] = red[0];
}

// out_k[e] = sum_j coef[k*m + j] X_j[e]   (k < kout <= 8)
template <class T>  // FP64 basis / sigma rows, or FP32 sigma rows (sums in FP64, rounded once)
__global__ __launch_bounds__(DL_BS) void k_lincomb(const T* __restrict__ X, int m, const double* __restrict__ coef,
                                                   int kout, int64_t n, T* __restrict__ out) {
    __shared__ double cf[DL_MAXM * 8];
    for (int i = threadIdx.x; i < m * kout; i += DL_BS) cf[i] = coef[i];
    __syncthreads();
    for (int64_t e = int64_t(blockIdx.x) * DL_BS + threadIdx.x; e < n; e += int64_t(gridDim.x) * DL_BS) {
        double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int j = 0; j < m; ++j) {
            double xv = double(X[size_t(j) * n + e]);
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (k < kout) acc[k] += cf[k * m + j] * xv;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (k < kout) out[size_t(k) * n + e] = T(acc[k]);
    }
}

// v -= sum_i d[i] A_i ; part[blk] = partial ||v||^2
__global__ __launch_bounds__(DL_BS) void k_gs_update(double* __restrict__ v, const double* __restrict__ A, int m,
                                                     const double* __restrict__ d, int64_t n, double* __restrict__ part) {
    __shared__ double dd[DL_MAXM];
    __shared__ double red[DL_BS / 64];
    for (int i = threadIdx.x; i < m; i += DL_BS) dd[i] = d[i];
    __syncthreads();
    double s2 = 0.0;
    for (int64_t e = int64_t(blockIdx.x) * DL_BS + threadIdx.x; e < n; e += int64_t(gridDim.x) * DL_BS) {
        double x = v[e];
        for (int i = 0; i < m; ++i) x -= dd[i] * A[size_t(i) * n + e];
        v[e] = x;
        s2 += x * x;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1) s2 += __shfl_down(s2, off, 64);
    if (lane == 0) red[wave] = s2;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < DL_BS / 64; ++w) s += red[w];
        part[blockIdx.x] = s;
    }
}

__global__ void k_scale(double* __restrict__ v, int64_t n, double a) {
    int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < n) v[e] *= a;
}

// residual r = sum_j a_j sigma_j - lambda sum_j a_j b_j; part = ||r||^2 partial; then the
// diagonal preconditioner r /= (lambda - h) with |denom| <= 1e-6 -> 0 (davidson_liu_solver.cc:412-448)
template <class TS>
__global__ __launch_bounds__(DL_BS) void k_residual(const TS* __restrict__ Sg, const double* __restrict__ B, int m,
                                                    const double* __restrict__ a, double lambda,
                                                    const double* __restrict__ h, int precond, int64_t n,
                                                    double* __restrict__ r, double* __restrict__ part) {
    __shared__ double ca[DL_MAXM];
    __shared__ double red[DL_BS / 64];
    for (int i = threadIdx.x; i < m; i += DL_BS) ca[i] = a[i];
    __syncthreads();
    double s2 = 0.0;
    for (int64_t e = int64_t(blockIdx.x) * DL_BS + threadIdx.x; e < n; e += int64_t(gridDim.x) * DL_BS) {
        double rs = 0.0, tb = 0.0;
        for (int j = 0; j < m; ++j) {
            rs += ca[j] * double(Sg[size_t(j) * n + e]);
            tb += ca[j] * B[size_t(j) * n + e];
        }
        double x = rs - lambda * tb;
        s2 += x * x;
        if (precond) {
            double den = lambda - h[e];
            x = (fabs(den) > 1.0e-6) ? x / den : 0.0;
        }
        r[e] = x;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1) s2 += __shfl_down(s2, off, 64);
    if (lane == 0) red[wave] = s2;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < DL_BS / 64; ++w) s += red[w];
        part[blockIdx.x] = s;
    }
}

// k_residual with the basis loop unrolled to MB (>= m) so that all 2m loads of an element are in
// flight together; the per-element arithmetic and its order are k_residual's. OV: the partials of
// the corrected vector's norm and of its overlaps with the basis vectors are accumulated in the
// same sweep (part stride m + 2: [0] ||r||^2 before preconditioning, [1] after, [2 + i] r . B_i).
template <int MB, bool OV, class TS>
__global__ __launch_bounds__(DL_BS) void k_residual_t(const TS* __restrict__ Sg, const double* __restrict__ B,
                                                      int m, const double* __restrict__ a, double lambda,
                                                      const double* __restrict__ h, int precond, int64_t n,
                                                      double* __restrict__ r, double* __restrict__ part) {
    constexpr int NA = OV ? MB + 2 : 1;
    __shared__ double ca[MB];
    __shared__ double red[NA][DL_BS / 64];
    for (int i = threadIdx.x; i < m; i += DL_BS) ca[i] = a[i];
    __syncthreads();
    double acc[NA];
#pragma unroll
    for (int i = 0; i < NA; ++i) acc[i] = 0.0;
    for (int64_t e = int64_t(blockIdx.x) * DL_BS + threadIdx.x; e < n; e += int64_t(gridDim.x) * DL_BS) {
        // coefficients re-read from LDS each element rather than held in 2m registers: more
        // waves per SIMD to hide the loads
        asm volatile("" ::: "memory");
        double sv[MB], bv[MB];
#pragma unroll
        for (int j = 0; j < MB; ++j)
            if (j < m) {
                sv[j] = double(Sg[size_t(j) * n + e]);
                bv[j] = B[size_t(j) * n + e];
            }
        double rs = 0.0, tb = 0.0;
#pragma unroll
        for (int j = 0; j < MB; ++j)
            if (j < m) {
                rs += ca[j] * sv[j];
                tb += ca[j] * bv[j];
            }
        double x = rs - lambda * tb;
        acc[0] += x * x;
        if (precond) {
            const double den = lambda - h[e];
            x = (fabs(den) > 1.0e-6) ? x / den : 0.0;
        }
        r[e] = x;
        if constexpr (OV) {
            acc[1] += x * x;
#pragma unroll
            for (int j = 0; j < MB; ++j)
                if (j < m) acc[2 + j] += x * bv[j];
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int na = OV ? m + 2 : 1;
    for (int i = 0; i < NA; ++i) {
        if (i >= na) break;
        double t = acc[i];
        for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
        if (lane == 0) red[i][wave] = t;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < na; i += DL_BS) {
        double t = 0.0;
        for (int w = 0; w < DL_BS / 64; ++w) t += red[i][w];
        part[size_t(blockIdx.x) * na + i] = t;
    }
}

// v <- a v and the partials of ||v||^2 of the result
__global__ __launch_bounds__(DL_BS) void k_scale_norm(double* __restrict__ v, int64_t n, double a,
                                                      double* __restrict__ part) {
    __shared__ double red[DL_BS / 64];
    double s2 = 0.0;
    for (int64_t e = int64_t(blockIdx.x) * DL_BS + threadIdx.x; e < n; e += int64_t(gridDim.x) * DL_BS) {
        const double x = a * v[e];
        v[e] = x;
        s2 += x * x;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1) s2 += __shfl_down(s2, off, 64);
    if (lane == 0) red[wave] = s2;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < DL_BS / 64; ++w) t += red[w];
        part[blockIdx.x] = t;
    }
}

// the diagonal preconditioner alone, r /= (lambda - h) with |denom| <= 1e-6 -> 0
// (davidson_liu_solver.cc:432-448), bit-identical to k_residual's precond branch
__global__ void k_diag_precond(double* __restrict__ r, const double* __restrict__ h, double lambda, int64_t n) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double den = lambda - h[e];
    r[e] = (fabs(den) > 1.0e-6) ? r[e] / den : 0.0;
}

// Olsen's correction (J. Olsen, P. Jorgensen, J. Simons, Chem. Phys. Lett. 169, 463 (1990)):
// delta = (lambda - D)^-1 r - eps (lambda - D)^-1 x with eps = x.(lambda - D)^-1 r / x.(lambda - D)^-1 x,
// the diagonal correction made orthogonal to the Ritz vector x in the metric of (lambda - D)^-1.
// Pass 1: partial sums of x.u (u = the already-preconditioned residual) and x.(x / den).
__global__ __launch_bounds__(256) void k_olsen_dots(const double* __restrict__ u, const double* __restrict__ x,
                                                    const double* __restrict__ h, double lambda, int64_t n,
                                                    double* __restrict__ part) {
    __shared__ double red[2][4];
    double a = 0.0, b = 0.0;
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n; e += int64_t(gridDim.x) * blockDim.x) {
        const double den = lambda - h[e];
        const double w = (fabs(den) > 1.0e-6) ? x[e] / den : 0.0;
        a += x[e] * u[e];
        b += x[e] * w;
    }
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off, 64);
        b += __shfl_down(b, off, 64);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][wave] = a;
        red[1][wave] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        part[2 * blockIdx.x + 1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    }
}
// pass 2: u -= eps x / den
__global__ void k_olsen_update(double* __restrict__ u, const double* __restrict__ x, const double* __restrict__ h,
                               double lambda, double eps, int64_t n) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double den = lambda - h[e];
    if (fabs(den) > 1.0e-6) u[e] -= eps * x[e] / den;
}

// project_out_roots (davidson_liu_solver.cc:590-603): sequential over bad roots, one block
__global__ void k_project_out(double* __restrict__ v, const int64_t* __restrict__ off, int nbad,
                              const int64_t* __restrict__ idx, const double* __restrict__ val) {
    __shared__ double red[256];
    for (int b = 0; b < nbad; ++b) {
        double s = 0.0;
        for (int64_t e = off[b] + threadIdx.x; e < off[b + 1]; e += blockDim.x) s += v[idx[e]] * val[e];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int st = blockDim.x / 2; st > 0; st >>= 1) {
            if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
            __syncthreads();
        }
        double ov = red[0];
        __syncthreads();
        // entries of one sparse vector have distinct indices
        for (int64_t e = off[b] + threadIdx.x; e < off[b + 1]; e += blockDim.x) v[idx[e]] -= ov * val[e];
        __syncthreads();
    }
}

__global__ void k_random(double* __restrict__ v, int64_t n, uint64_t seed) {
    int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n) return;
    uint64_t z = seed + 0x9E3779B97F4A7C15ULL * uint64_t(e + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    v[e] = 2.0 * ((double(z >> 11) + 0.5) * 0x1.0p-53) - 1.0;
}

// ---- DavidsonLiu -------------------------------------------------------------------------------
DavidsonLiu::DavidsonLiu(int64_t size, int nroot, int cpr, int spr) : size_(size), nroot_(nroot) {
    // startup(), davidson_liu_solver.cc:60-103
    if (size == 0) throw Error(-1, "DavidsonLiuSolver called with space of dimension zero.");
    if (nroot == 0) throw Error(-1, "DavidsonLiuSolver called with zero roots.");
    if (cpr < 1)
        throw Error(-1, "DavidsonLiuSolver: collapse_per_root_ (" + std::to_string(cpr) +
                            ") must be greater or equal to the number of roots (" + std::to_string(nroot) + ")");
    if (spr < cpr + 1)
        throw Error(-1, "DavidsonLiuSolver: subspace_per_root_ (" + std::to_string(spr) +
                            ") must be greater or equal to collapse_per_root_ + 1 (" + std::to_string(cpr + 1) + ")");
    collapse_size_ = int(std::min<int64_t>(int64_t(cpr) * nroot, size));
    subspace_size_ = int(std::min<int64_t>(int64_t(spr) * nroot, size));
    if (subspace_size_ > DL_MAXM) throw Error(-1, "DavidsonLiuSolver: subspace larger than 64 vectors");
    if (nroot_ > 8 && nroot_ > 0) {
        // rotations write at most 8 vectors per pass; larger counts are split in rotate()
    }
    size_t m = size_t(subspace_size_);
    b_.alloc(m * size_);
    sigma_v_.alloc(m * size_);
    temp_.alloc(m * size_);
    r_.alloc(m * size_);
    part_.alloc(size_t(DL_NBLK) * DL_MAXM + DL_MAXM);
    small_.alloc(size_t(DL_MAXM) * 8);
    HIPCHK(hipMemset(b_.p, 0, m * size_ * sizeof(double)));
    HIPCHK(hipMemset(sigma_v_.p, 0, m * size_ * sizeof(double)));
    lambda_.assign(m, 0.0);
    lambda_old_.assign(m, 0.0);
    resid_.assign(nroot_, 0.0);
    alpha_.assign(m * m, 0.0);
    G_.assign(m * m, 0.0);
    S_.assign(m * m, 0.0);
}

void DavidsonLiu::add_h_diag_host(const double* h, int64_t n) {
    if (n != size_)
        throw Error(-1, "DavidsonLiuSolver: h_diag vector size (" + std::to_string(n) +
                            ") must be equal to space size (" + std::to_string(size_) + ")");
    if (!hdiag_.p) hdiag_.alloc(size_);
    hdiag_.upload(h, n);
    have_hdiag_ = true;
}
void DavidsonLiu::add_h_diag_device(const double* d, int64_t n) {
    if (n != size_)
        throw Error(-1, "DavidsonLiuSolver: h_diag vector size (" + std::to_string(n) +
                            ") must be equal to space size (" + std::to_string(size_) + ")");
    if (!hdiag_.p) hdiag_.alloc(size_);
    HIPCHK(hipMemcpy(hdiag_.p, d, n * sizeof(double), hipMemcpyDeviceToDevice));
    have_hdiag_ = true;
}

void DavidsonLiu::add_sigma_builder32(Sigma32Fn f) {
    sigma32_ = std::move(f);
    s32_ = true;
    const size_t m = size_t(subspace_size_);
    if (sigma32_v_.n < m * size_t(size_)) {
        sigma32_v_.alloc(m * size_);
        HIPCHK(hipMemset(sigma32_v_.p, 0, m * size_ * sizeof(float)));
    }
}

void DavidsonLiu::add_project_out_vectors(const std::vector<SparseVec>& v) {
    poff_dev_.release();  // re-uploaded on the next projection
    poff_.assign(1, 0);
    std::vector<int64_t> idx;
    std::vector<double> val;
    for (const auto& x : v) {
        for (const auto& [i, c] : x) {
            if (i < 0 || i >= size_) throw Error(-1, "DavidsonLiuSolver: project-out index out of range");
            idx.push_back(i);
            val.push_back(c);
        }
        poff_.push_back(int64_t(idx.size()));
    }
    if (!idx.empty()) {
        pidx_.alloc(idx.size());
        pidx_.upload(idx.data(), idx.size());
        pval_.alloc(val.size());
        pval_.upload(val.data(), val.size());
    }
}

void DavidsonLiu::reset() {
    // davidson_liu_solver.cc:145-150
    basis_size_ = 0;
    sigma_size_ = 0;
    HIPCHK(hipMemset(b_.p, 0, b_.n * sizeof(double)));
    HIPCHK(hipMemset(sigma_v_.p, 0, sigma_v_.n * sizeof(double)));
}

void DavidsonLiu::eigenvector_host(int k, double* out) const {
    HIPCHK(hipMemcpy(out, b_.p + size_t(k) * size_, size_ * sizeof(double), hipMemcpyDeviceToHost));
}

template <class TX, class TY>
void DavidsonLiu::dots_t(const TX* x, const TY* Y, int m, double* out) {
    if (m <= 0) return;
    double* part = part_.p;
    double* res = part_.p + size_t(DL_NBLK) * DL_MAXM;
    int nb;
    // two-element loads need every vector aligned to two elements: even lengths only
    if (m <= DL_MAXM && size_ % 2 == 0 && reinterpret_cast<uintptr_t>(x) % (2 * sizeof(TX)) == 0 &&
        reinterpret_cast<uintptr_t>(Y) % (2 * sizeof(TY)) == 0) {
        nb = int(std::min<int64_t>(DL_NBLK, (size_ / 2 + DL_BS - 1) / DL_BS));
        nb = std::max(nb, 1);
        k_dots8<TX, TY><<<dim3(nb, (m + DOT_MC - 1) / DOT_MC), DL_BS>>>(x, Y, m, size_, part);
    } else {
        nb = int(std::min<int64_t>(DL_NBLK, (size_ + DL_BS - 1) / DL_BS));
        k_dots<TX, TY><<<nb, DL_BS>>>(x, Y, m, size_, part);
    }
    k_reduce_parts<<<m, 256>>>(part, nb, m, res);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(out, res, m * sizeof(double), hipMemcpyDeviceToHost));
}
void DavidsonLiu::dots(const double* x, const double* Y, int m, double* out) { dots_t(x, Y, m, out); }

// v <- a v - sum_i d_i A_i; returns ||v||^2 and ov[i] = v . A_i of the updated v
double DavidsonLiu::gs_fused(double* v, const double* src, const double* A, int m, const std::vector<double>& d,
                             double a, std::vector<double>& ov) {
    const int nb = int(std::min<int64_t>(DL_NBLK, (size_ + DL_BS - 1) / DL_BS));
    small_.upload(d.data(), std::max(m, 1));
    double* res = part_.p + size_t(DL_NBLK) * DL_MAXM;
    if (m <= 8)
        k_gs_fused<8><<<nb, DL_BS>>>(v, src, A, m, small_.p, a, size_, part_.p);
    else if (m <= 16)
        k_gs_fused<16><<<nb, DL_BS>>>(v, src, A, m, small_.p, a, size_, part_.p);
    else
        k_gs_fused<GS_MAXM><<<nb, DL_BS>>>(v, src, A, m, small_.p, a, size_, part_.p);
    k_reduce_parts<<<m + 1, 256>>>(part_.p, nb, m + 1, res);
    HIPCHK(hipGetLastError());
    std::vector<double> out(m + 1);
    HIPCHK(hipMemcpy(out.data(), res, (m + 1) * sizeof(double), hipMemcpyDeviceToHost));
    ov.assign(out.begin(), out.begin() + m);
    return out[m];
}

double DavidsonLiu::scale_norm(double* v, double a) {
    const int nb = int(std::min<int64_t>(DL_NBLK, (size_ + DL_BS - 1) / DL_BS));
    double* res = part_.p + size_t(DL_NBLK) * DL_MAXM;
    k_scale_norm<<<nb, DL_BS>>>(v, size_, a, part_.p);
    k_reduce_parts<<<1, 256>>>(part_.p, nb, 1, res);
    HIPCHK(hipGetLastError());
    double nn = 0.0;
    HIPCHK(hipMemcpy(&nn, res, sizeof(double), hipMemcpyDeviceToHost));
    return nn;
}

double DavidsonLiu::residual(int k, bool precond, std::vector<double>* ov) {
    const int m = subspace_size_, bs = basis_size_;
    std::vector<double> a(bs);
    for (int j = 0; j < bs; ++j) a[j] = alpha_[size_t(j) * m + k];
    ak_.upload(a.data(), bs);
    const int nb = int(std::min<int64_t>(DL_NBLK, (size_ + DL_BS - 1) / DL_BS));
    double* res = part_.p + size_t(DL_NBLK) * DL_MAXM;
    double* rk = r_.p + size_t(k) * size_;
    const int pc = precond ? 1 : 0;
    if (ov && bs > 16) throw Error(-1, "DavidsonLiu::residual: overlaps need a basis of at most 16 vectors");
    auto launch = [&](auto* Sg) {
        using TS = std::remove_const_t<std::remove_pointer_t<decltype(Sg)>>;
        if (bs <= 8 && ov)
            k_residual_t<8, true, TS><<<nb, DL_BS>>>(Sg, b_.p, bs, ak_.p, lambda_[k], hdiag_.p, pc, size_, rk, part_.p);
        else if (ov)
            k_residual_t<16, true, TS><<<nb, DL_BS>>>(Sg, b_.p, bs, ak_.p, lambda_[k], hdiag_.p, pc, size_, rk, part_.p);
        else if (bs <= 8)
            k_residual_t<8, false, TS><<<nb, DL_BS>>>(Sg, b_.p, bs, ak_.p, lambda_[k], hdiag_.p, pc, size_, rk, part_.p);
        else if (bs <= 16)
            k_residual_t<16, false, TS><<<nb, DL_BS>>>(Sg, b_.p, bs, ak_.p, lambda_[k], hdiag_.p, pc, size_, rk, part_.p);
        else  // larger bases: the loop kernel (same arithmetic, fewer registers)
            k_residual<TS><<<nb, DL_BS>>>(Sg, b_.p, bs, ak_.p, lambda_[k], hdiag_.p, pc, size_, rk, part_.p);
    };
    if (s32_)
        launch(static_cast<const float*>(sigma32_v_.p));
    else
        launch(static_cast<const double*>(sigma_v_.p));
    const int na = ov ? bs + 2 : 1;
    k_reduce_parts<<<na, 256>>>(part_.p, nb, na, res);
    HIPCHK(hipGetLastError());
    std::vector<double> out(na);
    HIPCHK(hipMemcpy(out.data(), res, na * sizeof(double), hipMemcpyDeviceToHost));
    if (ov) ov->assign(out.begin() + 1, out.end());
    return out[0];
}

template <class T>
void DavidsonLiu::lincomb_t(const T* X, int m, const double* coef, int k, T* out) {
    // coef: k outputs x m inputs (coef[kk*m + j]); split outputs in groups of 8
    int nb = int(std::min<int64_t>(4 * DL_NBLK, (size_ + DL_BS - 1) / DL_BS));
    for (int k0 = 0; k0 < k; k0 += 8) {
        int kk = std::min(8, k - k0);
        small_.upload(coef + size_t(k0) * m, size_t(kk) * m);  // synchronous copy
        k_lincomb<T><<<nb, DL_BS>>>(X, m, small_.p, kk, size_, out + size_t(k0) * size_);
        HIPCHK(hipGetLastError());
    }
}
void DavidsonLiu::lincomb(const double* X, int m, const double* coef, int k, double* out) {
    lincomb_t(X, m, coef, k, out);
}

// out_k = sum_j alpha[j][k] X_j for k < kout over the current basis (temp_->gemm(true,false,...))
template <class T>
void DavidsonLiu::rotate_t(const T* X, T* out, int kout) {
    int m = subspace_size_;
    std::vector<double> coef(size_t(kout) * basis_size_);
    for (int k = 0; k < kout; ++k)
        for (int j = 0; j < basis_size_; ++j) coef[size_t(k) * basis_size_ + j] = alpha_[size_t(j) * m + k];
    lincomb_t(X, basis_size_, coef.data(), kout, out);
}
void DavidsonLiu::rotate(const double* X, double* out, int kout) { rotate_t(X, out, kout); }

// Olsen's correction of root k's preconditioned residual (r_ row k), with its Ritz vector formed
// in a scratch row
void DavidsonLiu::olsen_correct(int k) {
    if (olsen_x_.n < size_t(size_)) olsen_x_.alloc(size_);
    const int m = subspace_size_;
    std::vector<double> coef(basis_size_);
    for (int j = 0; j < basis_size_; ++j) coef[j] = alpha_[size_t(j) * m + k];
    lincomb(b_.p, basis_size_, coef.data(), 1, olsen_x_.p);
    double* u = r_.p + size_t(k) * size_;
    const int nb = int(std::min<int64_t>(DL_NBLK, (size_ + 255) / 256));
    k_olsen_dots<<<nb, 256>>>(u, olsen_x_.p, hdiag_.p, lambda_[k], size_, part_.p);
    double* res = part_.p + size_t(DL_NBLK) * DL_MAXM;
    k_reduce_parts<<<2, 256>>>(part_.p, nb, 2, res);
    HIPCHK(hipGetLastError());
    double ab[2];
    HIPCHK(hipMemcpy(ab, res, 2 * sizeof(double), hipMemcpyDeviceToHost));
    if (std::fabs(ab[1]) < 1.0e-300) return;
    k_olsen_update<<<nblk(size_, 256), 256>>>(u, olsen_x_.p, hdiag_.p, lambda_[k], ab[0] / ab[1], size_);
    HIPCHK(hipGetLastError());
}

// Gram-Schmidt of vsrc against the rows of A into row rowsA, in fused passes: each cycle
// subtracts the projections found by the previous pass and, in the same sweep, measures the norm
// and the overlaps of the result; the normalization of a cycle is folded into the next update
// (v <- v / |v| - sum_i (ov_i / |v|) A_i). The first pass reads vsrc directly (no copy).
double DavidsonLiu::add_row_fused(double* A, int rowsA, const double* vsrc, bool* ok, double a0,
                                  const std::vector<double>* d0) {
    double* v = A + size_t(rowsA) * size_;
    std::vector<double> d(std::max(rowsA, 1));
    *ok = false;
    double a = 1.0;
    if (d0) {
        a = a0;
        for (int i = 0; i < rowsA; ++i) d[i] = (*d0)[i];
    } else if (rowsA > 0) {
        dots(vsrc, A, rowsA, d.data());
    }
    std::vector<double> ov;
    const double* src = vsrc;
    for (int cycle = 0; cycle < 10; ++cycle) {
        const double s2 = gs_fused(v, src, A, rowsA, d, a, ov);
        src = v;
        const double normval = std::sqrt(s2);
        if (normval < discard_) return 0.0;
        // after scaling by 1/normval: overlaps ov_i / normval, norm 1 up to rounding
        double mo = 0.0;
        for (int i = 0; i < rowsA; ++i) mo = std::max(mo, std::fabs(ov[i] / normval));
        if (mo < orth_) {
            const double nn = scale_norm(v, 1. / normval);
            if (std::fabs(nn - 1.0) >= orth_) {
                a = 1.0;
                std::fill(d.begin(), d.end(), 0.0);
                continue;
            }
            *ok = true;
            int m = subspace_size_;
            if (A == b_.p) {
                for (int i = 0; i < rowsA; ++i) S_[size_t(i) * m + rowsA] = S_[size_t(rowsA) * m + i] = ov[i] / normval;
                S_[size_t(rowsA) * m + rowsA] = nn;
            }
            return 1.0;
        }
        a = 1.0 / normval;
        for (int i = 0; i < rowsA; ++i) d[i] = ov[i] / normval;
    }
    return 0.0;
}

// add_row_and_orthonormalize (davidson_liu_solver.cc:648-687): returns norm check pass via ok.
// Gram-Schmidt against the rows of A is applied to all rows at once per cycle (classical GS);
// the cycle structure and thresholds are the reference's.
double DavidsonLiu::add_row_and_orthonormalize(double* A, int rowsA, const double* vsrc, bool* ok) {
    double* v = A + size_t(rowsA) * size_;
    if (rowsA <= GS_MAXM) return add_row_fused(A, rowsA, vsrc, ok);
    if (v != vsrc) HIPCHK(hipMemcpy(v, vsrc, size_ * sizeof(double), hipMemcpyDeviceToDevice));
    int nb = int(std::min<int64_t>(DL_NBLK, (size_ + DL_BS - 1) / DL_BS));
    std::vector<double> d(std::max(rowsA, 1)), parts(nb);
    *ok = false;
    for (int cycle = 0; cycle < 10; ++cycle) {
        if (rowsA > 0) dots(v, A, rowsA, d.data());
        small_.upload(d.data(), std::max(rowsA, 1));
        k_gs_update<<<nb, DL_BS>>>(v, A, rowsA, small_.p, size_, part_.p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(parts.data(), part_.p, nb * sizeof(double), hipMemcpyDeviceToHost));
        double s2 = 0.0;
        for (double x : parts) s2 += x;
        double normval = std::sqrt(s2);
        if (normval < discard_) return 0.0;
        k_scale<<<nblk(size_, 256), 256>>>(v, size_, 1. / normval);
        HIPCHK(hipGetLastError());
        std::vector<double> ov(rowsA + 1);
        // overlaps with the previous rows and the norm (A_rowsA is v itself)
        dots(v, A, rowsA + 1, ov.data());
        double mo = 0.0;
        for (int i = 0; i < rowsA; ++i) mo = std::max(mo, std::fabs(ov[i]));
        double nn = ov[rowsA];
        if (mo < orth_ && std::fabs(nn - 1.0) < orth_) {
            *ok = true;
            // cache the overlaps for check_orthonormality
            int m = subspace_size_;
            if (A == b_.p) {
                for (int i = 0; i < rowsA; ++i) S_[size_t(i) * m + rowsA] = S_[size_t(rowsA) * m + i] = ov[i];
                S_[size_t(rowsA) * m + rowsA] = nn;
            }
            return 1.0;
        }
    }
    return 0.0;
}

int DavidsonLiu::add_rows_and_orthonormalize(double* A, int rowsA, const double* B, int rowsB) {
    if (rowsA + rowsB > subspace_size_)
        throw Error(-5, "DavidsonLiuSolver: rowsA + rowsB (" + std::to_string(rowsA + rowsB) +
                            ") must be less or equal to matrix size (" + std::to_string(subspace_size_) + ")");
    int added = 0;
    for (int j = 0; j < rowsB; ++j) {
        bool ok = false;
        add_row_and_orthonormalize(A, rowsA + added, B + size_t(j) * size_, &ok);
        if (ok) ++added;
    }
    return added;
}

void DavidsonLiu::project_out(double* v, int nvec) {
    int nbad = int(poff_.size()) - 1;
    if (nbad <= 0) return;
    if (poff_dev_.n < poff_.size()) {
        poff_dev_.alloc(poff_.size());
        poff_dev_.upload(poff_.data(), poff_.size());
    }
    for (int k = 0; k < nvec; ++k)
        k_project_out<<<1, 256>>>(v + size_t(k) * size_, poff_dev_.p, nbad, pidx_.p, pval_.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(0));
}

__global__ void k_apply_mask(double* __restrict__ v, const uint8_t* __restrict__ m, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n && !m[i]) v[i] = 0.0;
}

void DavidsonLiu::fill_random(double* v, int nvec) {
    for (int k = 0; k < nvec; ++k) {
        rng_ += 0x9E3779B97F4A7C15ULL;
        k_random<<<nblk(size_, 256), 256>>>(v + size_t(k) * size_, size_, rng_);
        if (mask_) k_apply_mask<<<nblk(size_, 256), 256>>>(v + size_t(k) * size_, mask_, size_);
    }
    HIPCHK(hipGetLastError());
}

void DavidsonLiu::form_G() {
    // G = b sigma^T (davidson_liu_solver.cc:385-386). Entries between vectors that did not
    // change since the last call are kept; new columns and rows are computed.
    int m = subspace_size_;
    std::vector<double> tmp(m);
    int gv = std::min(gvalid_, basis_size_);
    for (int j = gv; j < basis_size_; ++j) {
        if (s32_)
            dots_t(sigma32_v_.p + size_t(j) * size_, b_.p, basis_size_, tmp.data());
        else
            dots(sigma_v_.p + size_t(j) * size_, b_.p, basis_size_, tmp.data());
        for (int i = 0; i < basis_size_; ++i) G_[size_t(i) * m + j] = tmp[i];
    }
    for (int i = gv; i < basis_size_; ++i) {
        if (gv == 0) break;
        if (s32_)
            dots_t(b_.p + size_t(i) * size_, static_cast<const float*>(sigma32_v_.p), gv, tmp.data());
        else
            dots(b_.p + size_t(i) * size_, sigma_v_.p, gv, tmp.data());
        for (int j = 0; j < gv; ++j) G_[size_t(i) * m + j] = tmp[j];
    }
    gvalid_ = basis_size_;
}

void DavidsonLiu::check_orthonormality() {
    // davidson_liu_solver.cc:689-730 (overlaps cached from the Gram-Schmidt passes)
    int m = subspace_size_;
    double thr = orth_ * 3.0, md = 0.0, mo = 0.0;
    for (int i = 0; i < basis_size_; ++i)
        for (int j = 0; j < basis_size_; ++j) {
            double s = S_[size_t(i) * m + j];
            if (i == j)
                md = std::max(md, std::fabs(s - 1.0));
            else
                mo = std::max(mo, std::fabs(s));
        }
    if (md > thr) throw Error(-5, "DavidsonLiuSolver::check_orthogonality(): eigenvectors are not normalized!");
    if (mo > thr) throw Error(-5, "DavidsonLiuSolver::check_orthogonality(): eigenvectors are not orthogonal!");
}

void DavidsonLiu::get_results() {
    // davidson_liu_solver.cc:488-501
    lambda_old_ = lambda_;
    rotate(b_.p, temp_.p, nroot_);
    HIPCHK(hipMemset(b_.p, 0, b_.n * sizeof(double)));
    std::fill(S_.begin(), S_.end(), 0.0);
    basis_size_ = 0;
    int added = add_rows_and_orthonormalize(b_.p, 0, temp_.p, nroot_);
    if (added != nroot_)
        throw Error(-5, "DavidsonLiuSolver: get_results generated less vectors (" + std::to_string(added) +
                            ") than expected (" + std::to_string(nroot_) + ")");
    basis_size_ = nroot_;
    sigma_size_ = 0;  // sigma rows no longer match the rotated basis
    gvalid_ = 0;
}

bool DavidsonLiu::solve() {
    const int m = subspace_size_;
    iters_ = 0;
    nsigma_ = 0;
    // setup_guesses, davidson_liu_solver.cc:276-330
    if (basis_size_ == 0) {
        int ng = int(guesses_.size());
        if (ng >= nroot_ && ng <= m) {
            HIPCHK(hipMemset(temp_.p, 0, temp_.n * sizeof(double)));
            for (int g = 0; g < ng; ++g) {
                std::vector<int64_t> idx;
                std::vector<double> val;
                for (const auto& [i, c] : guesses_[g]) {
                    if (i < 0 || i >= size_) throw Error(-1, "DavidsonLiuSolver: guess index out of range");
                    idx.push_back(i);
                    val.push_back(c);
                }
                // set_vector: M->set(k, I, CI) (later duplicates overwrite)
                std::vector<double> host_row;
                for (size_t t = 0; t < idx.size(); ++t)
                    HIPCHK(hipMemcpy(temp_.p + size_t(g) * size_ + idx[t], &val[t], sizeof(double), hipMemcpyHostToDevice));
            }
        } else if (ng == 0) {
            HIPCHK(hipMemset(temp_.p, 0, temp_.n * sizeof(double)));
            fill_random(temp_.p, nroot_);
        } else {
            throw Error(-1, "DavidsonLiuSolver: number of guess vectors (" + std::to_string(ng) +
                                ") must be between 0 and the subspace size (" + std::to_string(m) + ")");
        }
        int should = std::max(nroot_, ng);
        project_out(temp_.p, nroot_);
        std::fill(S_.begin(), S_.end(), 0.0);
        int added = add_rows_and_orthonormalize(b_.p, 0, temp_.p, should);
        if (added != should) throw Error(-5, "DavidsonLiuSolver: guess vectors are zero or linearly dependent");
        basis_size_ = added;
        sigma_size_ = 0;
    } else if (basis_size_ >= nroot_) {
        basis_size_ = nroot_;
        sigma_size_ = 0;
    } else {
        throw Error(-1, "DavidsonLiuSolver: number of guess vectors must be between 0 and the subspace size");
    }
    // preiteration_sanity_checks, :332-345
    if (s32_ ? !sigma32_ : !sigma_) throw Error(-1, "DavidsonLiuSolver: sigma builder has not been set");
    if (!have_hdiag_) throw Error(-1, "DavidsonLiuSolver: h_diag has not been set");

    std::fill(lambda_old_.begin(), lambda_old_.end(), 0.0);
    gvalid_ = 0;
    const int nb = int(std::min<int64_t>(DL_NBLK, (size_ + DL_BS - 1) / DL_BS));
    std::vector<double> parts(nb);
    for (int iter = 0; iter < maxiter_; ++iter) {
        iters_ = iter + 1;
        check_orthonormality();
        // compute_sigma, :374-383
        for (int j = sigma_size_; j < basis_size_; ++j) {
            if (s32_)
                sigma32_(b_.p + size_t(j) * size_, sigma32_v_.p + size_t(j) * size_);
            else
                sigma_(b_.p + size_t(j) * size_, sigma_v_.p + size_t(j) * size_);
            ++nsigma_;
        }
        sigma_size_ = basis_size_;
        // form_and_diagonalize_effective_hamiltonian, :385-410
        form_G();
        int bs = basis_size_;
        std::vector<double> Gb(size_t(bs) * bs), ev, V;
        for (int i = 0; i < bs; ++i)
            for (int j = 0; j < bs; ++j) Gb[size_t(i) * bs + j] = 0.5 * (G_[size_t(i) * m + j] + G_[size_t(j) * m + i]);
        jacobi_eigh(bs, Gb, ev, V);
        std::fill(alpha_.begin(), alpha_.end(), 0.0);
        std::fill(lambda_.begin(), lambda_.end(), 0.0);
        for (int k = 0; k < bs; ++k) {
            lambda_[k] = ev[k];
            for (int i = 0; i < bs; ++i) alpha_[size_t(i) * m + k] = V[size_t(k) * bs + i];
        }
        // form_residual_vectors + compute_residual_norm + form_correction_vectors, :412-455
        if (ak_.n < size_t(m)) ak_.alloc(m);
        // One root, the diagonal preconditioner, nothing to project out and no collapse ahead:
        // the residual sweep also measures the correction's norm and its overlaps with the basis,
        // which replaces the normalization pass and the first overlap pass of Gram-Schmidt.
        const bool fuse = nroot_ == 1 && !precond_ && !olsen_ && poff_.size() <= 1 && basis_size_ + nroot_ <= m &&
                          basis_size_ <= 16;  // the overlaps stay in registers up to 16 vectors
        std::vector<double> fov;
        for (int k = 0; k < nroot_; ++k)
            resid_[k] = std::sqrt(residual(k, !precond_, fuse ? &fov : nullptr));
        if (olsen_ && !precond_)
            for (int k = 0; k < nroot_; ++k) olsen_correct(k);
        if (precond_) {
            if (!precond_(r_.p, lambda_.data(), nroot_, iter)) {
                // the custom preconditioner declined: the diagonal one, in place on the residuals
                for (int k = 0; k < nroot_; ++k)
                    k_diag_precond<<<nblk(size_, 256), 256>>>(r_.p + size_t(k) * size_, hdiag_.p, lambda_[k], size_);
                HIPCHK(hipGetLastError());
            }
        }
        // project_out_roots(r_), normalize_vectors(r_, nroot_), :457-465
        project_out(r_.p, nroot_);
        if (!fuse) {
            for (int k = 0; k < nroot_; ++k) {
                double nr2 = 0.0;
                dots(r_.p + size_t(k) * size_, r_.p + size_t(k) * size_, 1, &nr2);
                k_scale<<<nblk(size_, 256), 256>>>(r_.p + size_t(k) * size_, size_, 1.0 / std::sqrt(nr2));
            }
            HIPCHK(hipGetLastError());
        }
        if (print_ >= 2) {
            double avg = 0.0, mx_de = 0.0, mx_r = 0.0;
            for (int k = 0; k < nroot_; ++k) {
                avg += lambda_[k] / nroot_;
                mx_de = std::max(mx_de, std::fabs(lambda_[k] - lambda_old_[k]));
                mx_r = std::max(mx_r, resid_[k]);
            }
            printf("  %6d  %20.12f  %20.12f  %20.12f %6d\n", iter, avg, mx_de, mx_r, basis_size_);
            fflush(stdout);
        }
        // check_convergence, :467-486
        int ne = 0, nr = 0;
        for (int k = 0; k < nroot_; ++k) {
            ne += std::fabs(lambda_[k] - lambda_old_[k]) < e_conv_;
            nr += resid_[k] < r_conv_;
        }
        bool e_ok = ne == nroot_, r_ok = nr == nroot_;
        bool stop = observe_ && observe_(iter, resid_, lambda_);
        if ((e_ok && r_ok) || int64_t(basis_size_) == size_ || stop) {
            get_results();
            return true;
        }
        // subspace_collapse, :538-588
        if (basis_size_ + nroot_ > m) {
            int cs = std::min(std::max(collapse_size_, nroot_), basis_size_);
            rotate(b_.p, temp_.p, cs);
            // rows past the basis size are never read, so b_ is not cleared (13 GB at CAS(16,16))
            std::fill(S_.begin(), S_.end(), 0.0);
            int keep_bs = basis_size_;
            basis_size_ = 0;
            int added = add_rows_and_orthonormalize(b_.p, 0, temp_.p, cs);
            basis_size_ = keep_bs;
            if (s32_) {  // temp_ holds cs FP32 rows in its first half
                float* t32 = reinterpret_cast<float*>(temp_.p);
                rotate_t(static_cast<const float*>(sigma32_v_.p), t32, cs);
                HIPCHK(hipMemcpy(sigma32_v_.p, t32, size_t(cs) * size_ * sizeof(float), hipMemcpyDeviceToDevice));
            } else {
                rotate(sigma_v_.p, temp_.p, cs);
                HIPCHK(hipMemcpy(sigma_v_.p, temp_.p, size_t(cs) * size_ * sizeof(double), hipMemcpyDeviceToDevice));
            }
            if (added != cs)
                throw Error(-5, "DavidsonLiuSolver: collapse_vectors generated less vectors (" + std::to_string(added) +
                                    ") than expected (" + std::to_string(cs) + ")");
            basis_size_ = cs;
            sigma_size_ = cs;
            gvalid_ = 0;
        }
        // add the correction vectors, :221-251
        int num_to_add = std::min(nroot_, m - basis_size_);
        int added = 0;
        if (fuse && basis_size_ + 1 <= GS_MAXM) {
            // normalized correction r / |r| with overlaps (r . b_i) / |r|
            const double a0 = 1.0 / std::sqrt(fov[0]);
            std::vector<double> d0(basis_size_);
            for (int i = 0; i < basis_size_; ++i) d0[i] = fov[1 + i] * a0;
            bool ok = false;
            add_row_fused(b_.p, basis_size_, r_.p, &ok, a0, &d0);
            added = ok ? 1 : 0;
        } else {
            if (fuse) {  // the normalization skipped above
                double nr2 = fov[0];
                k_scale<<<nblk(size_, 256), 256>>>(r_.p, size_, 1.0 / std::sqrt(nr2));
                HIPCHK(hipGetLastError());
            }
            added = add_rows_and_orthonormalize(b_.p, basis_size_, r_.p, num_to_add);
        }
        basis_size_ += added;
        int missing = num_to_add - added;
        if (missing > 0) {
            HIPCHK(hipMemset(temp_.p, 0, size_t(missing) * size_ * sizeof(double)));
            fill_random(temp_.p, missing);
            project_out(temp_.p, std::min(missing, nroot_));
            int ra = add_rows_and_orthonormalize(b_.p, basis_size_, temp_.p, missing);
            basis_size_ += ra;
            added += ra;
        }
        if (added == 0) {
            get_results();
            return r_ok;
        }
        lambda_old_ = lambda_;
    }
    get_results();
    return false;
}

}  // namespace fci
