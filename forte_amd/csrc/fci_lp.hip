// fci_lp.hip -- learned Davidson preconditioner for the gfx950 FCI solver.
//
// Forte's Davidson-Liu solver corrects a residual with the diagonal preconditioner
// d_I = r_I / (lambda - H_II), elements with |lambda - H_II| <= 1e-6 set to zero
// (forte/helpers/davidson_liu_solver.cc:432-448). The learned variant multiplies each element by
// exp(y_I), y_I = MLP(x_I), so y = 0 is the reference preconditioner. The MLP is
// 34 -> 128 -> 128 -> 1 with the tanh form of GELU after the first two layers; its features are
//   x_I = [(H_II - lambda) / 16, r_I / rms(r), 16 alpha bits, 16 beta bits]
// with the bits describing each spin string relative to its aufbau reference: bit j < 8 is 1 when
// orbital k-1-j is empty, bit 8+j is 1 when orbital k+j is occupied (0 outside 0..n-1).
//
// Parameters (FP32, this order): W1[128][34], b1[128], W2[128][128], b2[128], w3[128], b3[1].
//
// Inference (k_lp_apply) runs on the BF16 MFMA with FP32 accumulation: weights in BF16, the 32
// occupation bits exact, the two continuous features and b1 as BF16 hi+lo pairs. Training runs in
// FP32 on the vector ALUs (tiled GEMM + element-wise kernels + Adam).
#include "fci_lp.h"

#include <cmath>
#include <cstring>

namespace fci {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

static inline unsigned nblk(int64_t n, int bs) { return unsigned((n + bs - 1) / bs); }

constexpr float kGeluC = 0.7978845608028654f;  // sqrt(2 / pi)
constexpr float kGeluA = 0.044715f;

// GELU, tanh form (Hendrycks and Gimpel, 2016); tanh(u) = 1 - 2 / (1 + exp(2u))
__device__ __forceinline__ float gelu_t(float z, float* tanh_out) {
    const float u = kGeluC * (z + kGeluA * z * z * z);
    const float t = 1.f - 2.f / (1.f + __expf(2.f * u));
    if (tanh_out) *tanh_out = t;
    return 0.5f * z * (1.f + t);
}
__device__ __forceinline__ float gelu_grad(float z) {
    float t;
    gelu_t(z, &t);
    return 0.5f * (1.f + t) + 0.5f * z * (1.f - t * t) * kGeluC * (1.f + 3.f * kGeluA * z * z);
}

// hole/particle bits 0..7 (holes: orbital k-1-j empty) or 8..15 (particles: orbital k+j occupied)
__device__ __forceinline__ float lp_bit(uint32_t s, int n, int k, int j) {
    if (j < 8) {
        const int o = k - 1 - j;
        return (o >= 0 && o < n) ? float(((s >> o) & 1u) ^ 1u) : 0.f;
    }
    const int o = k + (j - 8);
    return (o >= 0 && o < n) ? float((s >> o) & 1u) : 0.f;
}

// ---- inference ----------------------------------------------------------------------------------
// Packed BF16 weights, one 8-element vector per (step, tile, lane) in MFMA A-operand order:
//   step 0: W1[16t + l%16][2 + 8(l/16) + x]                       (occupation bits)
//   step 1: lanes l/16 == 0: [W1[h][0], W1[h][0], W1[h][1], W1[h][1], b1_hi[h], b1_lo[h], 0, 0]
//   step 2+ks: W2[16t + l%16][(2ks + x/4) * 16 + 4 (l/16) + x%4]  (the k order that matches the
//              layer-1 accumulator layout, so layer-2 B operands need no data exchange)
constexpr int LP_NSTEP = 6;
constexpr int LP_WPK = LP_NSTEP * 8 * 64;  // bf8 vectors

__global__ __launch_bounds__(256) void k_lp_apply(double* __restrict__ r, const double* __restrict__ h, double lam,
                                                  double inv_rms, int64_t N, int64_t Nb,
                                                  const uint32_t* __restrict__ str_a,
                                                  const uint32_t* __restrict__ str_b, int na, int nb, int n,
                                                  const bf8* __restrict__ wpk, const float* __restrict__ wf) {
    __shared__ bf8 Ws[LP_WPK];          // 48 KiB
    __shared__ float Wf[2 * LP_NH + 1];  // b2, w3, b3
    for (int i = threadIdx.x; i < LP_WPK; i += 256) Ws[i] = wpk[i];
    for (int i = threadIdx.x; i < 2 * LP_NH + 1; i += 256) Wf[i] = wf[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, cl = lane & 15;
    const int64_t ntile = (N + 15) / 16;
    // two tiles of 16 samples per pass: each weight vector read from LDS feeds both, and the two
    // independent MFMA chains overlap
    const int64_t stride = int64_t(gridDim.x) * 4;
    for (int64_t tile0 = int64_t(blockIdx.x) * 4 + wave; tile0 < ntile; tile0 += 2 * stride) {
        // the weights are re-read from LDS every pass rather than held in ~200 registers: that
        // keeps several waves per SIMD resident to hide the latency of the per-sample loads
        asm volatile("" ::: "memory");
        double rv[2], hv[2];
        bool ok[2];
        int64_t Iv[2];
        bf8 bits[2], cont[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int64_t I = (tile0 + q * stride) * 16 + cl;
            Iv[q] = I;
            ok[q] = I < N;
            rv[q] = ok[q] ? r[I] : 0.0;
            hv[q] = ok[q] ? h[I] : 0.0;
            // B operand of layer 1: bits 8(g&1)..8(g&1)+7 of the alpha (g < 2) or beta string
            const int64_t Ia = ok[q] ? I / Nb : 0, Ib = ok[q] ? I % Nb : 0;
            const uint32_t st = g < 2 ? str_a[Ia] : str_b[Ib];
            const int k = g < 2 ? na : nb;
#pragma unroll
            for (int x = 0; x < 8; ++x) {
                bits[q][x] = __bf16(ok[q] ? lp_bit(st, n, k, 8 * (g & 1) + x) : 0.f);
                cont[q][x] = __bf16(0.f);
            }
            if (g == 0 && ok[q]) {
                const float f0 = float((hv[q] - lam) * (1.0 / 16.0));
                const float f1 = float(rv[q] * inv_rms);
                cont[q][0] = __bf16(f0);
                cont[q][1] = __bf16(f0 - float(cont[q][0]));
                cont[q][2] = __bf16(f1);
                cont[q][3] = __bf16(f1 - float(cont[q][2]));
                cont[q][4] = __bf16(1.f);
                cont[q][5] = __bf16(1.f);
            }
        }
        f4 z[2][8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const bf8 w0 = Ws[(0 * 8 + t) * 64 + lane], w1 = Ws[(1 * 8 + t) * 64 + lane];
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                z[q][t] = f4{0.f, 0.f, 0.f, 0.f};
                z[q][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w0, bits[q], z[q][t], 0, 0, 0);
                z[q][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1, cont[q], z[q][t], 0, 0, 0);
            }
        }
        // z[q][t][v] = Z1[hidden 16t + 4g + v][sample cl of tile q]
        bf8 b2[2][4];
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int ks = 0; ks < 4; ++ks)
#pragma unroll
                for (int x = 0; x < 8; ++x) b2[q][ks][x] = __bf16(gelu_t(z[q][2 * ks + x / 4][x % 4], nullptr));
        float part[2] = {0.f, 0.f};
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            f4 u[2] = {f4{0.f, 0.f, 0.f, 0.f}, f4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                const bf8 w = Ws[((2 + ks) * 8 + t) * 64 + lane];
#pragma unroll
                for (int q = 0; q < 2; ++q) u[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, b2[q][ks], u[q], 0, 0, 0);
            }
#pragma unroll
            for (int q = 0; q < 2; ++q)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int h2 = 16 * t + 4 * g + v;
                    part[q] += Wf[LP_NH + h2] * gelu_t(u[q][v] + Wf[h2], nullptr);
                }
        }
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            float pq = part[q];
            pq += __shfl_xor(pq, 16, 64);
            pq += __shfl_xor(pq, 32, 64);
            float y = pq + Wf[2 * LP_NH];
            y = fminf(fmaxf(y, -30.f), 30.f);
            if (g == 0 && ok[q]) {
                const double den = lam - hv[q];
                r[Iv[q]] = (fabs(den) > 1.0e-6) ? rv[q] / den * double(__expf(y)) : 0.0;
            }
        }
    }
}

static std::vector<__bf16> pack_inference_weights(const float* p) {
    const float* W1 = p + LP_OFF_W1;
    const float* b1 = p + LP_OFF_B1;
    const float* W2 = p + LP_OFF_W2;
    std::vector<__bf16> w(size_t(LP_WPK) * 8, __bf16(0.f));
    auto at = [&](int step, int t, int l, int x) -> __bf16& { return w[((size_t(step) * 8 + t) * 64 + l) * 8 + x]; };
    for (int t = 0; t < 8; ++t)
        for (int l = 0; l < 64; ++l) {
            const int hr = 16 * t + (l & 15), g = l >> 4;
            for (int x = 0; x < 8; ++x) at(0, t, l, x) = __bf16(W1[hr * LP_NIN + 2 + 8 * g + x]);
            if (g == 0) {
                at(1, t, l, 0) = __bf16(W1[hr * LP_NIN + 0]);
                at(1, t, l, 1) = __bf16(W1[hr * LP_NIN + 0]);
                at(1, t, l, 2) = __bf16(W1[hr * LP_NIN + 1]);
                at(1, t, l, 3) = __bf16(W1[hr * LP_NIN + 1]);
                const __bf16 hi = __bf16(b1[hr]);
                at(1, t, l, 4) = hi;
                at(1, t, l, 5) = __bf16(b1[hr] - float(hi));
            }
            for (int ks = 0; ks < 4; ++ks)
                for (int x = 0; x < 8; ++x) {
                    const int h1 = (2 * ks + x / 4) * 16 + 4 * g + (x % 4);
                    at(2 + ks, t, l, x) = __bf16(W2[hr * LP_NH + h1]);
                }
        }
    return w;
}

void LearnedPrecond::set_weights(const float* p) {
    params_.assign(p, p + LP_NPARAM);
    for (float x : params_)
        if (!std::isfinite(x)) throw Error(-1, "learned preconditioner weights are not finite");
    auto w = pack_inference_weights(p);
    static_assert(sizeof(__bf16) == sizeof(uint16_t), "bf16 storage");
    wpk_.alloc(w.size());
    HIPCHK(hipMemcpy(wpk_.p, w.data(), w.size() * sizeof(__bf16), hipMemcpyHostToDevice));
    std::vector<float> f(2 * LP_NH + 1);
    std::memcpy(f.data(), p + LP_OFF_B2, LP_NH * sizeof(float));
    std::memcpy(f.data() + LP_NH, p + LP_OFF_W3, LP_NH * sizeof(float));
    f[2 * LP_NH] = p[LP_OFF_B3];
    wf_.alloc(f.size());
    wf_.upload(f.data(), f.size());
}

void LearnedPrecond::apply(double* r, const double* hdiag, double lambda, double rnorm, const SpinTables& a,
                           const SpinTables& b, hipStream_t st) const {
    if (!ready()) throw Error(-2, "learned preconditioner weights have not been set");
    const int64_t N = a.nstr * b.nstr;
    const double rms = rnorm / std::sqrt(double(N));
    const double inv = rms > 0.0 ? 1.0 / rms : 0.0;
    const int64_t ntile = (N + 15) / 16;
    const unsigned grid = unsigned(std::min<int64_t>((ntile + 3) / 4, 256 * 8));
    k_lp_apply<<<grid, 256, 0, st>>>(r, hdiag, lambda, inv, N, b.nstr, a.str.p, b.str.p, a.k, b.k, a.n,
                                     reinterpret_cast<const bf8*>(wpk_.p), wf_.p);
    HIPCHK(hipGetLastError());
}

// ---- features of sampled determinants (training data) ------------------------------------------
__global__ void k_lp_features(const int64_t* __restrict__ idx, int64_t m, const double* __restrict__ h, double lam,
                              const double* __restrict__ r, double inv_rms, int64_t Nb,
                              const uint32_t* __restrict__ str_a, const uint32_t* __restrict__ str_b, int na,
                              int nb, int n, float* __restrict__ X) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int64_t I = idx[i];
    const uint32_t sa = str_a[I / Nb], sb = str_b[I % Nb];
    float* x = X + i * LP_NIN;
    x[0] = float((h[I] - lam) * (1.0 / 16.0));
    x[1] = float(r[I] * inv_rms);
    for (int j = 0; j < 16; ++j) {
        x[2 + j] = lp_bit(sa, n, na, j);
        x[18 + j] = lp_bit(sb, n, nb, j);
    }
}

void lp_features_device(const SpinTables& a, const SpinTables& b, const int64_t* d_idx, int64_t m,
                        const double* d_h, double lambda, const double* d_r, double rnorm, float* d_X,
                        hipStream_t st) {
    const int64_t N = a.nstr * b.nstr;
    const double rms = rnorm / std::sqrt(double(N));
    k_lp_features<<<nblk(m, 256), 256, 0, st>>>(d_idx, m, d_h, lambda, d_r, rms > 0 ? 1.0 / rms : 0.0, b.nstr,
                                                a.str.p, b.str.p, a.k, b.k, a.n, d_X);
    HIPCHK(hipGetLastError());
}

// d0[i] = r / (lambda - h) at idx[i] (the diagonal correction, 0 where |lambda - h| <= 1e-6) and
// xs[i] = x[idx[i]] (the Ritz vector the residual belongs to)
__global__ void k_lp_sample(const int64_t* __restrict__ idx, int64_t m, const double* __restrict__ h, double lam,
                            const double* __restrict__ r, const double* __restrict__ x, float* __restrict__ d0,
                            double* __restrict__ xs) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int64_t I = idx[i];
    const double den = lam - h[I];
    d0[i] = float(fabs(den) > 1.0e-6 ? r[I] / den : 0.0);
    xs[i] = x[I];
}

// e[i] = c[idx[i]] / cx - xs[i]: the exact correction c / (c.x) - x at the sampled determinants
__global__ void k_lp_target(const int64_t* __restrict__ idx, int64_t m, const double* __restrict__ c, double inv_cx,
                            const double* __restrict__ xs, float* __restrict__ e) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= m) return;
    e[i] = float(c[idx[i]] * inv_cx - xs[i]);
}

__global__ __launch_bounds__(256) void k_dot_parts(const double* __restrict__ a, const double* __restrict__ b,
                                                   int64_t n, double* __restrict__ part) {
    __shared__ double red[4];
    double s = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) s += a[i] * b[i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

double dot_device(const double* a, const double* b, int64_t n, hipStream_t st) {
    const int nb = int(std::min<int64_t>((n + 255) / 256, 1024));
    DBuf<double> part(nb);
    k_dot_parts<<<nb, 256, 0, st>>>(a, b, n, part.p);
    HIPCHK(hipGetLastError());
    std::vector<double> h(nb);
    HIPCHK(hipStreamSynchronize(st));
    part.download(h.data(), nb);
    double t = 0.0;
    for (double x : h) t += x;
    return t;
}

void lp_sample_device(const int64_t* d_idx, int64_t m, const double* d_h, double lambda, const double* d_r,
                      const double* d_x, float* d_d0, double* d_xs, hipStream_t st) {
    k_lp_sample<<<nblk(m, 256), 256, 0, st>>>(d_idx, m, d_h, lambda, d_r, d_x, d_d0, d_xs);
    HIPCHK(hipGetLastError());
}
void lp_target_device(const int64_t* d_idx, int64_t m, const double* d_c, double cx, const double* d_xs,
                      float* d_e, hipStream_t st) {
    k_lp_target<<<nblk(m, 256), 256, 0, st>>>(d_idx, m, d_c, 1.0 / cx, d_xs, d_e);
    HIPCHK(hipGetLastError());
}

// ---- training (FP32) ----------------------------------------------------------------------------
// C[M][N] = alpha * op(A) op(B) + beta * C, op(A)[m][k] = TA ? A[k*lda + m] : A[m*lda + k],
// op(B)[k][n] = TB ? B[n*ldb + k] : B[k*ldb + n]. 64x64 tiles, 16-deep k slices, 4x4 per thread.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm_f32(int M, int N, int K, float alpha, const float* __restrict__ A,
                                                  int lda, const float* __restrict__ B, int ldb, float beta,
                                                  float* __restrict__ C, int ldc) {
    __shared__ float As[16][65];
    __shared__ float Bs[16][65];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    float acc[4][4] = {};
    for (int k0 = 0; k0 < K; k0 += 16) {
        for (int i = threadIdx.x; i < 16 * 64; i += 256) {
            const int kk = i / 64, mm = i % 64;
            const int gm = m0 + mm, gk = k0 + kk;
            As[kk][mm] = (gm < M && gk < K) ? (TA ? A[size_t(gk) * lda + gm] : A[size_t(gm) * lda + gk]) : 0.f;
            const int gn = n0 + mm;
            Bs[kk][mm] = (gn < N && gk < K) ? (TB ? B[size_t(gn) * ldb + gk] : B[size_t(gk) * ldb + gn]) : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            float a[4], bb[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) bb[j] = Bs[kk][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * bb[j];
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gm = m0 + ty * 4 + i, gn = n0 + tx * 4 + j;
            if (gm < M && gn < N) {
                float* c = C + size_t(gm) * ldc + gn;
                *c = alpha * acc[i][j] + (beta != 0.f ? beta * *c : 0.f);
            }
        }
}

template <bool TA, bool TB>
static void gemm(int M, int N, int K, float alpha, const float* A, int lda, const float* B, int ldb, float beta,
                 float* C, int ldc, hipStream_t st) {
    dim3 grid((N + 63) / 64, (M + 63) / 64);
    k_gemm_f32<TA, TB><<<grid, 256, 0, st>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
    HIPCHK(hipGetLastError());
}

// Z[i][j] += b[j]; A = gelu(Z)
__global__ void k_bias_gelu(float* __restrict__ Z, float* __restrict__ Aout, const float* __restrict__ b, int64_t m,
                            int nh) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= m * nh) return;
    const float z = Z[e] + b[e % nh];
    Z[e] = z;
    Aout[e] = gelu_t(z, nullptr);
}

// D[i][j] = (gy[i] * w[j] or D[i][j]) * gelu'(Z[i][j])
__global__ void k_gelu_back(float* __restrict__ D, const float* __restrict__ Z, const float* __restrict__ gy,
                            const float* __restrict__ w, int64_t m, int nh) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= m * nh) return;
    const float up = gy ? gy[e / nh] * w[e % nh] : D[e];
    D[e] = up * gelu_grad(Z[e]);
}

// out[j] = sum_i D[i][j] (fixed order)
__global__ void k_colsum(const float* __restrict__ D, int64_t m, int nh, float* __restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nh) return;
    float s = 0.f;
    for (int64_t i = 0; i < m; ++i) s += D[i * nh + j];
    out[j] = s;
}

// loss of one group and dloss/dy; y holds MLP(x) + b3 on entry. One block.
__global__ __launch_bounds__(1024) void k_lp_loss(const float* __restrict__ y, const float* __restrict__ d0,
                                                  const float* __restrict__ e, int64_t m, float reg,
                                                  float* __restrict__ gy, float* __restrict__ loss_out,
                                                  float* __restrict__ gb3) {
    __shared__ float red[4][16];
    __shared__ float tot[4];
    float s[4] = {0.f, 0.f, 0.f, 0.f};  // de, dd, ee, yy
    for (int64_t i = threadIdx.x; i < m; i += 1024) {
        const float yc = fminf(fmaxf(y[i], -30.f), 30.f);
        const float d = d0[i] * __expf(yc);
        s[0] += d * e[i];
        s[1] += d * d;
        s[2] += e[i] * e[i];
        s[3] += y[i] * y[i];
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        for (int off = 32; off > 0; off >>= 1) s[q] += __shfl_down(s[q], off, 64);
        if (lane == 0) red[q][wave] = s[q];
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        float t = 0.f;
        for (int w = 0; w < 16; ++w) t += red[threadIdx.x][w];
        tot[threadIdx.x] = t;
    }
    __syncthreads();
    const float Sde = tot[0], Sdd = tot[1], See = tot[2];
    // a group with a vanishing correction or target carries no direction: no loss, no gradient
    const bool ok = Sdd > 1e-30f && See > 1e-30f && isfinite(Sdd * See) && isfinite(Sde);
    const float inv = ok ? 1.f / (Sdd * See) : 0.f;
    if (threadIdx.x == 0) *loss_out = ok ? 1.f - Sde * Sde * inv + reg * tot[3] / float(m) : 0.f;
    float sg = 0.f;
    for (int64_t i = ok ? int64_t(threadIdx.x) : m; i < m; i += 1024) {
        const float yc = fminf(fmaxf(y[i], -30.f), 30.f);
        const float d = d0[i] * __expf(yc);
        const float gd = -2.f * Sde * e[i] * inv + 2.f * Sde * Sde * d * inv / Sdd;
        const float gyi = gd * d + 2.f * reg * y[i] / float(m);
        gy[i] = gyi;
        sg += gyi;
    }
    if (!ok)
        for (int64_t i = threadIdx.x; i < m; i += 1024) gy[i] = 0.f;
    for (int off = 32; off > 0; off >>= 1) sg += __shfl_down(sg, off, 64);
    __syncthreads();
    if (lane == 0) red[0][wave] = sg;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.f;
        for (int w = 0; w < 16; ++w) t += red[0][w];
        *gb3 = t;
    }
}

// Adam (Kingma and Ba, 2015) with bias correction
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float v = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = v;
    p[i] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

// y += b3 (no activation on the output layer)
__global__ void k_add_b3(float* __restrict__ y, const float* __restrict__ b3, int64_t m) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < m) y[i] += *b3;
}

LPTrainer::LPTrainer(int64_t group) : G_(group) {
    if (group < 1) throw Error(-1, "group size must be positive");
    p_.alloc(LP_NPARAM);
    g_.alloc(LP_NPARAM);
    m1_.alloc(LP_NPARAM);
    m2_.alloc(LP_NPARAM);
    z1_.alloc(size_t(G_) * LP_NH);
    a1_.alloc(size_t(G_) * LP_NH);
    z2_.alloc(size_t(G_) * LP_NH);
    a2_.alloc(size_t(G_) * LP_NH);
    y_.alloc(G_);
    gy_.alloc(G_);
    loss_.alloc(1);
}

void LPTrainer::set_params(const float* p) {
    p_.upload(p, LP_NPARAM);
    HIPCHK(hipMemset(m1_.p, 0, LP_NPARAM * sizeof(float)));
    HIPCHK(hipMemset(m2_.p, 0, LP_NPARAM * sizeof(float)));
    t_ = 0;
}
void LPTrainer::get_params(float* p) const { p_.download(p, LP_NPARAM); }

void LPTrainer::forward(const float* X, int64_t m, hipStream_t st) {
    if (m > G_) throw Error(-1, "batch larger than the trainer's group size");
    const float* P = p_.p;
    gemm<false, true>(int(m), LP_NH, LP_NIN, 1.f, X, LP_NIN, P + LP_OFF_W1, LP_NIN, 0.f, z1_.p, LP_NH, st);
    k_bias_gelu<<<nblk(m * LP_NH, 256), 256, 0, st>>>(z1_.p, a1_.p, P + LP_OFF_B1, m, LP_NH);
    gemm<false, true>(int(m), LP_NH, LP_NH, 1.f, a1_.p, LP_NH, P + LP_OFF_W2, LP_NH, 0.f, z2_.p, LP_NH, st);
    k_bias_gelu<<<nblk(m * LP_NH, 256), 256, 0, st>>>(z2_.p, a2_.p, P + LP_OFF_B2, m, LP_NH);
    gemm<false, false>(int(m), 1, LP_NH, 1.f, a2_.p, LP_NH, P + LP_OFF_W3, 1, 0.f, y_.p, 1, st);
    k_add_b3<<<nblk(m, 256), 256, 0, st>>>(y_.p, P + LP_OFF_B3, m);
    HIPCHK(hipGetLastError());
}

float LPTrainer::loss_grad(const float* X, const float* d0, const float* e, int64_t m, float reg, hipStream_t st) {
    forward(X, m, st);
    const float* P = p_.p;
    float* Gr = g_.p;
    k_lp_loss<<<1, 1024, 0, st>>>(y_.p, d0, e, m, reg, gy_.p, loss_.p, Gr + LP_OFF_B3);
    // output layer: gw3 = A2^T gy
    gemm<true, false>(LP_NH, 1, int(m), 1.f, a2_.p, LP_NH, gy_.p, 1, 0.f, Gr + LP_OFF_W3, 1, st);
    // D2 = (gy w3^T) o gelu'(Z2) (in a2_), gW2 = D2^T A1, gb2 = colsum D2
    k_gelu_back<<<nblk(m * LP_NH, 256), 256, 0, st>>>(a2_.p, z2_.p, gy_.p, P + LP_OFF_W3, m, LP_NH);
    gemm<true, false>(LP_NH, LP_NH, int(m), 1.f, a2_.p, LP_NH, a1_.p, LP_NH, 0.f, Gr + LP_OFF_W2, LP_NH, st);
    k_colsum<<<1, LP_NH, 0, st>>>(a2_.p, m, LP_NH, Gr + LP_OFF_B2);
    // D1 = (D2 W2) o gelu'(Z1) (in a1_), gW1 = D1^T X, gb1 = colsum D1
    gemm<false, false>(int(m), LP_NH, LP_NH, 1.f, a2_.p, LP_NH, P + LP_OFF_W2, LP_NH, 0.f, a1_.p, LP_NH, st);
    k_gelu_back<<<nblk(m * LP_NH, 256), 256, 0, st>>>(a1_.p, z1_.p, nullptr, nullptr, m, LP_NH);
    gemm<true, false>(LP_NH, LP_NIN, int(m), 1.f, a1_.p, LP_NH, X, LP_NIN, 0.f, Gr + LP_OFF_W1, LP_NIN, st);
    k_colsum<<<1, LP_NH, 0, st>>>(a1_.p, m, LP_NH, Gr + LP_OFF_B1);
    HIPCHK(hipGetLastError());
    float l = 0.f;
    HIPCHK(hipMemcpyAsync(&l, loss_.p, sizeof(float), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return l;
}

void LPTrainer::get_grad(float* g) const { g_.download(g, LP_NPARAM); }
void LPTrainer::get_y(float* y, int64_t m) const { y_.download(y, size_t(m)); }

void LPTrainer::adam(float lr, hipStream_t st) {
    ++t_;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, float(t_)), c2 = 1.f - std::pow(b2, float(t_));
    k_adam<<<nblk(LP_NPARAM, 256), 256, 0, st>>>(p_.p, g_.p, m1_.p, m2_.p, LP_NPARAM, lr, b1, b2, 1e-8f, c1, c2);
    HIPCHK(hipGetLastError());
}

}  // namespace fci
