// fci_sigma.hip -- sigma = H C on gfx950 (FCIVector::Hamiltonian, fci_vector_hamiltonian.cc:46-272).
//
// C is the N_a x N_b matrix of CI coefficients (alpha rows). The terms:
//
// Same spin (H1 + H2_aaaa / H2_bbbb, :93-192). For k >= 2 electrons of one spin the one-body
// operator is folded into the two-body one, sum_pq h_pq a+p a_q = 1/(k-1) sum_pqt h_pq
// a+p a+t a_t a_q, so the whole same-spin operator is sum_{P>Q,R>S} g'[PQ,RS] a+P a+Q a_S a_R.
// Inserting the identity over the (k-2)-electron hole strings K turns it into dense blocks:
// every K owns the M = C(n-k+2,2) strings I = K + {a,b}, and
//     Y[I_m, :] += s_m sum_m' g'[m,m'] s_m' X[I_m', :]          (one M x M GEMM per K)
// with s_m the sign of a+a a+b |K>. For k == 1 the hole string is empty and g' = h.
// The alpha side runs on the rows of C, the beta side on the rows of C^T (transposes).
//
// Alpha-beta (H2_aabb, :194-272). sigma[Ja,Jb] = sum (pq|rs) <Ja|E^a_pq|Ia><Jb|E^b_rs|Ib>
// C[Ia,Ib]. One block per alpha row Ja: gather X[e,Ib] = s_e C[Ia_e,Ib] over the k(n-k)
// single excitations into Ja plus the merged diagonal, contract on MFMA with
// V[rs,e] = (p_e q_e|rs) over packed pairs rs (Y = V X), and scatter Y[rs,Ib] to
// sigma[Ja, E^b_rs Ib] with LDS atomics into a row buffer.
//
// FP64: v_mfma_f64_16x16x4_f64. Mixed: FP32 storage, BF16x3 split products
// (hi*hi + hi*lo + lo*hi) on v_mfma_f32_16x16x32_bf16, FP32 accumulation.
#include "fci_internal.h"

#include <algorithm>
#include <cstring>

namespace fci {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

static inline unsigned nblk(int64_t n, int bs) { return unsigned((n + bs - 1) / bs); }

// Hardware dispatches workgroup b to XCD (b % 8); give each XCD a contiguous range of work.
__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t nwork) {
    int64_t per = (nwork + 7) / 8;
    return (b % 8) * per + b / 8;
}
static inline unsigned xcd_grid(int64_t nwork) { return unsigned(((nwork + 7) / 8) * 8); }

// ---- elementwise helpers ---------------------------------------------------------------------
template <class T> __global__ void k_fill(T* p, int64_t n, T v) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
template <class T> __global__ void k_scale_copy(const T* __restrict__ x, T* __restrict__ y, int64_t n, T a) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i];
}

// ---- transposes (32x32 tiles through LDS) ------------------------------------------------------
template <class T>
__global__ void k_transpose(const T* __restrict__ in, T* __restrict__ out, int64_t R, int64_t Cc,
                            const T* __restrict__ base, T bs) {
    __shared__ T tile[32][33];
    int64_t c0 = int64_t(blockIdx.x) * 32, r0 = int64_t(blockIdx.y) * 32;
    for (int j = threadIdx.y; j < 32; j += 8) {
        int64_t r = r0 + j, c = c0 + threadIdx.x;
        if (r < R && c < Cc) tile[j][threadIdx.x] = in[r * Cc + c];
    }
    __syncthreads();
    for (int j = threadIdx.y; j < 32; j += 8) {
        int64_t c = c0 + j, r = r0 + threadIdx.x;
        if (r < R && c < Cc) {
            T v = tile[threadIdx.x][j];
            if (base) v += bs * base[c * R + r];
            out[c * R + r] = v;
        }
    }
}

template <class T> void transpose_device(const T* in, T* out, int64_t rows, int64_t cols, hipStream_t st) {
    dim3 g(nblk(cols, 32), nblk(rows, 32)), b(32, 8);
    k_transpose<T><<<g, b, 0, st>>>(in, out, rows, cols, nullptr, T(0));
    HIPCHK(hipGetLastError());
}
template <class T>
void transpose_add_device(const T* in, T* out, int64_t rows, int64_t cols, const T* base, T bs,
                          hipStream_t st) {
    dim3 g(nblk(cols, 32), nblk(rows, 32)), b(32, 8);
    k_transpose<T><<<g, b, 0, st>>>(in, out, rows, cols, base, bs);
    HIPCHK(hipGetLastError());
}
template void transpose_device<double>(const double*, double*, int64_t, int64_t, hipStream_t);
template void transpose_device<float>(const float*, float*, int64_t, int64_t, hipStream_t);
template void transpose_add_device<double>(const double*, double*, int64_t, int64_t, const double*, double, hipStream_t);
template void transpose_add_device<float>(const float*, float*, int64_t, int64_t, const float*, float, hipStream_t);

// ---- same-spin pair-clique GEMM ----------------------------------------------------------------
constexpr int CQ_TW = 64;   // columns per block (4 waves x 16)
constexpr int CQ_MT = 48;   // output members per block (3 MFMA row tiles)
constexpr int CQ_MAXM = 512;

// FP64: E = W D with W = g'[members] (M x M), D = signed rows; Y[I_m] += s_m E[m].
__global__ __launch_bounds__(256) void k_clique_f64(const double* __restrict__ X, double* __restrict__ Y,
                                                    int64_t L, const uint32_t* __restrict__ clq, int M,
                                                    int64_t nK, const double* __restrict__ G, int gdim,
                                                    int nct, int nmt) {
    constexpr int KC = 16;
    __shared__ uint32_t mem[CQ_MAXM];
    __shared__ double Ws[CQ_MT][KC + 1];
    __shared__ double Ds[KC][CQ_TW];
    const int64_t nwork = int64_t(nct) * nK * nmt;
    const int64_t w = xcd_remap(blockIdx.x, nwork);
    if (w >= nwork) return;
    const int mt = int(w % nmt);
    const int64_t rest = w / nmt;
    const int64_t K = rest % nK;
    const int ct = int(rest / nK);
    const int64_t c0 = int64_t(ct) * CQ_TW;
    const int r0 = mt * CQ_MT;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < M; i += 256) mem[i] = clq[K * M + i];
    __syncthreads();
    d4 acc[3];
    for (int j = 0; j < 3; ++j) acc[j] = d4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < M; k0 += KC) {
        for (int i = tid; i < CQ_MT * KC; i += 256) {
            int r = i / KC, kk = i % KC;
            int gr = r0 + r, gk = k0 + kk;
            double v = 0.0;
            if (gr < M && gk < M) {
                int ir = int((mem[gr] >> 21) & 0x3FFu), ik = int((mem[gk] >> 21) & 0x3FFu);
                v = G[size_t(ir) * gdim + ik];
            }
            Ws[r][kk] = v;
        }
        for (int i = tid; i < KC * CQ_TW; i += 256) {
            int kk = i / CQ_TW, c = i % CQ_TW;
            int gk = k0 + kk;
            double v = 0.0;
            if (gk < M && c0 + c < L) {
                uint32_t e = mem[gk];
                v = X[int64_t(e & kAddrMask) * L + c0 + c];
                if (e >> 31) v = -v;
            }
            Ds[kk][c] = v;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < KC / 4; ++ks) {
            double b = Ds[ks * 4 + (lane >> 4)][wave * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                double a = Ws[j * 16 + (lane & 15)][ks * 4 + (lane >> 4)];
                acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    const int64_t col = c0 + wave * 16 + (lane & 15);
    if (col >= L) return;
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            int r = r0 + j * 16 + 4 * v + (lane >> 4);  // f64 16x16x4 output row
            if (r < M) {
                uint32_t e = mem[r];
                double val = (e >> 31) ? -acc[j][v] : acc[j][v];
                atomicAdd(&Y[int64_t(e & kAddrMask) * L + col], val);
            }
        }
}

__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo) {
    hi = __bf16(x);
    lo = __bf16(x - float(hi));
}

// Mixed: FP32 storage, BF16x3 products, FP32 accumulation.
__global__ __launch_bounds__(256) void k_clique_mixed(const float* __restrict__ X, float* __restrict__ Y,
                                                      int64_t L, const uint32_t* __restrict__ clq, int M,
                                                      int64_t nK, const float* __restrict__ G, int gdim,
                                                      int nct, int nmt) {
    constexpr int KC = 32;
    constexpr int KP = KC + 8;  // padded k stride (bf16 elements)
    __shared__ uint32_t mem[CQ_MAXM];
    __shared__ __attribute__((aligned(16))) __bf16 Wh[CQ_MT][KP], Wl[CQ_MT][KP];
    __shared__ __attribute__((aligned(16))) __bf16 Dh[CQ_TW][KP], Dl[CQ_TW][KP];
    const int64_t nwork = int64_t(nct) * nK * nmt;
    const int64_t w = xcd_remap(blockIdx.x, nwork);
    if (w >= nwork) return;
    const int mt = int(w % nmt);
    const int64_t rest = w / nmt;
    const int64_t K = rest % nK;
    const int ct = int(rest / nK);
    const int64_t c0 = int64_t(ct) * CQ_TW;
    const int r0 = mt * CQ_MT;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < M; i += 256) mem[i] = clq[K * M + i];
    __syncthreads();
    f4 acc[3];
    for (int j = 0; j < 3; ++j) acc[j] = f4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < M; k0 += KC) {
        for (int i = tid; i < CQ_MT * KC; i += 256) {
            int r = i / KC, kk = i % KC;
            int gr = r0 + r, gk = k0 + kk;
            float v = 0.f;
            if (gr < M && gk < M) {
                int ir = int((mem[gr] >> 21) & 0x3FFu), ik = int((mem[gk] >> 21) & 0x3FFu);
                v = G[size_t(ir) * gdim + ik];
            }
            split_bf16(v, Wh[r][kk], Wl[r][kk]);
        }
        for (int i = tid; i < KC * CQ_TW; i += 256) {
            int kk = i / CQ_TW, c = i % CQ_TW;
            int gk = k0 + kk;
            float v = 0.f;
            if (gk < M && c0 + c < L) {
                uint32_t e = mem[gk];
                v = X[int64_t(e & kAddrMask) * L + c0 + c];
                if (e >> 31) v = -v;
            }
            split_bf16(v, Dh[c][kk], Dl[c][kk]);
        }
        __syncthreads();
        const int kb = 8 * (lane >> 4);
        bf8 bh = *reinterpret_cast<const bf8*>(&Dh[wave * 16 + (lane & 15)][kb]);
        bf8 bl = *reinterpret_cast<const bf8*>(&Dl[wave * 16 + (lane & 15)][kb]);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            bf8 ah = *reinterpret_cast<const bf8*>(&Wh[j * 16 + (lane & 15)][kb]);
            bf8 al = *reinterpret_cast<const bf8*>(&Wl[j * 16 + (lane & 15)][kb]);
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc[j], 0, 0, 0);
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc[j], 0, 0, 0);
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc[j], 0, 0, 0);
        }
        __syncthreads();
    }
    const int64_t col = c0 + wave * 16 + (lane & 15);
    if (col >= L) return;
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            int r = r0 + j * 16 + 4 * (lane >> 4) + v;  // f32-output MFMA row
            if (r < M) {
                uint32_t e = mem[r];
                float val = (e >> 31) ? -acc[j][v] : acc[j][v];
                atomicAdd(&Y[int64_t(e & kAddrMask) * L + col], val);
            }
        }
}

// Column-tile-major variant: the whole operator g' lives in LDS, each wave owns one hole string
// K at a time, and the XCD-contiguous work order keeps one column tile of X and Y resident in
// that XCD's L2 while every K of the tile passes through it. The MFMA A operand is read from
// g' through the member pair indices, so no per-K copy of W_K is made.
constexpr int CQ2_KCH = 64;    // hole strings per block

template <int TW, int CQ2_WAVES>
__global__ __launch_bounds__(64 * CQ2_WAVES) void k_clique2_f64(const double* __restrict__ X, double* __restrict__ Y,
                                                                int64_t L, const uint32_t* __restrict__ clq, int M,
                                                                int64_t nK, const double* __restrict__ G, int gdim,
                                                                int nct, int nkc) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    constexpr int NT = TW / 16;
    double* Gs = reinterpret_cast<double*>(smem);                     // [gdim][gdim]
    const int Mp = (M + 15) & ~15;
    double* Dall = Gs + size_t(gdim) * gdim;                          // per wave [Mp][TW]
    uint32_t* memall = reinterpret_cast<uint32_t*>(Dall + size_t(CQ2_WAVES) * Mp * TW);  // per wave [Mp]
    const int64_t nwork = int64_t(nct) * nkc;
    const int64_t w = xcd_remap(blockIdx.x, nwork);
    if (w >= nwork) return;
    const int kc = int(w % nkc), ct = int(w / nkc);
    const int64_t c0 = int64_t(ct) * TW;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < gdim * gdim; i += 64 * CQ2_WAVES) Gs[i] = G[i];
    __syncthreads();
    double* Ds = Dall + size_t(wave) * Mp * TW;
    uint32_t* mem = memall + size_t(wave) * Mp;
    const int nmt = Mp / 16;
    for (int64_t K = int64_t(kc) * CQ2_KCH + wave; K < min(int64_t(kc + 1) * CQ2_KCH, nK); K += CQ2_WAVES) {
        for (int i = lane; i < Mp; i += 64) mem[i] = i < M ? clq[K * M + i] : 0u;
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        for (int i = lane; i < Mp * TW; i += 64) {
            int m = i / TW, c = i % TW;
            double v = 0.0;
            if (m < M && c0 + c < L) {
                uint32_t e = mem[m];
                v = X[int64_t(e & kAddrMask) * L + c0 + c];
                if (e >> 31) v = -v;
            }
            Ds[i] = v;
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        for (int mt = 0; mt < nmt; ++mt) {
            d4 acc[NT];
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) acc[nt] = d4{0.0, 0.0, 0.0, 0.0};
            const int row = mt * 16 + (lane & 15);
            const int prow = row < M ? int((mem[row] >> 21) & 0x3FFu) : 0;
            for (int k0 = 0; k0 < Mp; k0 += 4) {
                const int kk = k0 + (lane >> 4);
                double a = 0.0;
                if (row < M && kk < M) a = Gs[prow * gdim + int((mem[kk] >> 21) & 0x3FFu)];
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) {
                    double b = Ds[kk * TW + nt * 16 + (lane & 15)];
                    acc[nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[nt], 0, 0, 0);
                }
            }
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const int64_t col = c0 + nt * 16 + (lane & 15);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = mt * 16 + 4 * v + (lane >> 4);  // f64 output row
                    if (r < M && col < L) {
                        uint32_t e = mem[r];
                        atomicAdd(&Y[int64_t(e & kAddrMask) * L + col], (e >> 31) ? -acc[nt][v] : acc[nt][v]);
                    }
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

template <int TW, int CQ2_WAVES>
__global__ __launch_bounds__(64 * CQ2_WAVES) void k_clique2_mixed(const float* __restrict__ X, float* __restrict__ Y,
                                                                  int64_t L, const uint32_t* __restrict__ clq, int M,
                                                                  int64_t nK, const float* __restrict__ G, int gdim,
                                                                  int nct, int nkc) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    constexpr int NT = TW / 16;
    float* Gs = reinterpret_cast<float*>(smem);                       // [gdim][gdim]
    const int Mk = (M + 31) & ~31;                                    // K padded to 32
    const int Mp = (M + 15) & ~15;                                    // rows padded to 16
    const int KPP = Mk + 8;
    __bf16* Dall = reinterpret_cast<__bf16*>(Gs + ((size_t(gdim) * gdim + 3) & ~size_t(3)));  // per wave hi,lo [TW][KPP]
    uint32_t* memall = reinterpret_cast<uint32_t*>(Dall + size_t(CQ2_WAVES) * 2 * TW * KPP);
    const int64_t nwork = int64_t(nct) * nkc;
    const int64_t w = xcd_remap(blockIdx.x, nwork);
    if (w >= nwork) return;
    const int kc = int(w % nkc), ct = int(w / nkc);
    const int64_t c0 = int64_t(ct) * TW;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < gdim * gdim; i += 64 * CQ2_WAVES) Gs[i] = G[i];
    __syncthreads();
    __bf16* Dh = Dall + size_t(wave) * 2 * TW * KPP;
    __bf16* Dl = Dh + TW * KPP;
    uint32_t* mem = memall + size_t(wave) * Mk;
    const int nmt = Mp / 16;
    for (int64_t K = int64_t(kc) * CQ2_KCH + wave; K < min(int64_t(kc + 1) * CQ2_KCH, nK); K += CQ2_WAVES) {
        for (int i = lane; i < Mk; i += 64) mem[i] = i < M ? clq[K * M + i] : 0u;
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        for (int i = lane; i < Mk * TW; i += 64) {
            int m = i / TW, c = i % TW;
            float v = 0.f;
            if (m < M && c0 + c < L) {
                uint32_t e = mem[m];
                v = X[int64_t(e & kAddrMask) * L + c0 + c];
                if (e >> 31) v = -v;
            }
            split_bf16(v, Dh[c * KPP + m], Dl[c * KPP + m]);
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        for (int mt = 0; mt < nmt; ++mt) {
            f4 acc[NT];
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) acc[nt] = f4{0.f, 0.f, 0.f, 0.f};
            const int row = mt * 16 + (lane & 15);
            const int prow = row < M ? int((mem[row] >> 21) & 0x3FFu) : 0;
            for (int k0 = 0; k0 < Mk; k0 += 32) {
                bf8 ah, al;
#pragma unroll
                for (int x = 0; x < 8; ++x) {
                    const int kk = k0 + 8 * (lane >> 4) + x;
                    float g = 0.f;
                    if (row < M && kk < M) g = Gs[prow * gdim + int((mem[kk] >> 21) & 0x3FFu)];
                    __bf16 h, l;
                    split_bf16(g, h, l);
                    ah[x] = h;
                    al[x] = l;
                }
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) {
                    const int off = (nt * 16 + (lane & 15)) * KPP + k0 + 8 * (lane >> 4);
                    bf8 bh = *reinterpret_cast<const bf8*>(Dh + off);
                    bf8 bl = *reinterpret_cast<const bf8*>(Dl + off);
                    acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc[nt], 0, 0, 0);
                    acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc[nt], 0, 0, 0);
                    acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc[nt], 0, 0, 0);
                }
            }
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const int64_t col = c0 + nt * 16 + (lane & 15);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = mt * 16 + 4 * (lane >> 4) + v;  // f32-output row
                    if (r < M && col < L) {
                        uint32_t e = mem[r];
                        atomicAdd(&Y[int64_t(e & kAddrMask) * L + col], (e >> 31) ? -acc[nt][v] : acc[nt][v]);
                    }
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// ---- alpha-beta: per alpha row gather -> MFMA -> LDS scatter ----------------------------------
constexpr int AB_TW = 64;      // Ib columns per tile (4 MFMA column tiles)
constexpr int AB_MAXE = 160;   // max single excitations + 1 per string
constexpr int AB_MTW = 3;      // max MFMA row tiles (of packed rs pairs) per wave

// KS: number of 4-wide k-steps (E padded to 4*KS).
template <int KS>
__global__ __launch_bounds__(256) void k_ab_f64(const double* __restrict__ C, double* __restrict__ S,
                                                int64_t Na, int64_t Nb, const uint32_t* __restrict__ into_a,
                                                int ns, const uint32_t* __restrict__ str_a,
                                                const uint32_t* __restrict__ out_b, int np,
                                                const double* __restrict__ vpk, int nmt, int use_lds_row) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    constexpr int KP = KS * 4;
    double* Xs = reinterpret_cast<double*>(smem);       // [KP][AB_TW]
    double* srow = Xs + KP * AB_TW;                     // [Nb] when use_lds_row
    __shared__ uint32_t ent[AB_MAXE];
    const int64_t Ja = xcd_remap(blockIdx.x, Na);
    if (Ja >= Na) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int E = ns + 1;
    for (int i = tid; i < ns; i += 256) ent[i] = into_a[Ja * ns + i];
    if (use_lds_row)
        for (int64_t i = tid; i < Nb; i += 256) srow[i] = 0.0;
    __syncthreads();
    const uint32_t occ = str_a[Ja];
    // A operand (V) in registers for this wave's row tiles: V[rs][e] = (p_e q_e | rs)
    double va[AB_MTW][KS];
#pragma unroll
    for (int j = 0; j < AB_MTW; ++j) {
        const int mtile = wave + 4 * j;
        const int rs = mtile * 16 + (lane & 15);
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const int e = ks * 4 + (lane >> 4);
            double v = 0.0;
            if (mtile < nmt && rs < np) {
                if (e < ns) {
                    v = vpk[size_t((ent[e] >> 21) & 0x3FFu) * np + rs];
                } else if (e == ns) {
                    uint32_t o = occ;
                    while (o) {
                        int q = __ffs(o) - 1;
                        o &= o - 1u;
                        v += vpk[size_t(q * (q + 1) / 2 + q) * np + rs];
                    }
                }
            }
            va[j][ks] = v;
        }
    }
    for (int64_t c0 = 0; c0 < Nb; c0 += AB_TW) {
        // scatter targets of this tile, loaded before the gather so their latency overlaps it
        uint32_t tgt[AB_MTW][4][4];
#pragma unroll
        for (int j = 0; j < AB_MTW; ++j)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int rs = (wave + 4 * j) * 16 + 4 * v + (lane >> 4);  // f64 output row
                    const int64_t Ib = c0 + nt * 16 + (lane & 15);
                    tgt[j][nt][v] = (rs < np && Ib < Nb) ? out_b[int64_t(rs) * Nb + Ib] : kInvalid;
                }
        for (int i = tid; i < KP * AB_TW; i += 256) {
            int e = i / AB_TW, c = i % AB_TW;
            double v = 0.0;
            if (c0 + c < Nb) {
                if (e < ns) {
                    uint32_t t = ent[e];
                    v = C[int64_t(t & kAddrMask) * Nb + c0 + c];
                    if (t >> 31) v = -v;
                } else if (e == ns) {
                    v = C[Ja * Nb + c0 + c];
                }
            }
            Xs[e * AB_TW + c] = v;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < AB_MTW; ++j) {
            const int mtile = wave + 4 * j;
            if (mtile >= nmt) break;
            d4 acc[4];
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) acc[nt] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
#pragma unroll
                for (int nt = 0; nt < 4; ++nt) {
                    double b = Xs[(ks * 4 + (lane >> 4)) * AB_TW + nt * 16 + (lane & 15)];
                    acc[nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(va[j][ks], b, acc[nt], 0, 0, 0);
                }
            }
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const uint32_t t = tgt[j][nt][v];
                    if (t == kInvalid) continue;
                    double y = (t >> 31) ? -acc[nt][v] : acc[nt][v];
                    if (use_lds_row)
                        atomicAdd(&srow[t & kAddrMask], y);
                    else
                        atomicAdd(&S[Ja * Nb + (t & kAddrMask)], y);
                }
            }
        }
        __syncthreads();
    }
    if (use_lds_row)
        for (int64_t i = tid; i < Nb; i += 256) S[Ja * Nb + i] += srow[i];
}

// KS: number of 32-wide k-steps.
template <int KS>
__global__ __launch_bounds__(256) void k_ab_mixed(const float* __restrict__ C, float* __restrict__ S,
                                                  int64_t Na, int64_t Nb, const uint32_t* __restrict__ into_a,
                                                  int ns, const uint32_t* __restrict__ str_a,
                                                  const uint32_t* __restrict__ out_b, int np,
                                                  const float* __restrict__ vpk, int nmt, int use_lds_row) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    constexpr int KP = KS * 32;
    constexpr int KPP = KP + 8;  // padded k stride
    __bf16* Xh = reinterpret_cast<__bf16*>(smem);     // [AB_TW][KPP]
    __bf16* Xl = Xh + AB_TW * KPP;
    float* srow = reinterpret_cast<float*>(Xl + AB_TW * KPP);
    __shared__ uint32_t ent[AB_MAXE];
    const int64_t Ja = xcd_remap(blockIdx.x, Na);
    if (Ja >= Na) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < ns; i += 256) ent[i] = into_a[Ja * ns + i];
    if (use_lds_row)
        for (int64_t i = tid; i < Nb; i += 256) srow[i] = 0.f;
    __syncthreads();
    const uint32_t occ = str_a[Ja];
    bf8 vh[AB_MTW][KS], vl[AB_MTW][KS];
#pragma unroll
    for (int j = 0; j < AB_MTW; ++j) {
        const int mtile = wave + 4 * j;
        const int rs = mtile * 16 + (lane & 15);
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
#pragma unroll
            for (int x = 0; x < 8; ++x) {
                const int e = ks * 32 + 8 * (lane >> 4) + x;
                float v = 0.f;
                if (mtile < nmt && rs < np) {
                    if (e < ns) {
                        v = vpk[size_t((ent[e] >> 21) & 0x3FFu) * np + rs];
                    } else if (e == ns) {
                        uint32_t o = occ;
                        while (o) {
                            int q = __ffs(o) - 1;
                            o &= o - 1u;
                            v += vpk[size_t(q * (q + 1) / 2 + q) * np + rs];
                        }
                    }
                }
                __bf16 h, l;
                split_bf16(v, h, l);
                vh[j][ks][x] = h;
                vl[j][ks][x] = l;
            }
        }
    }
    for (int64_t c0 = 0; c0 < Nb; c0 += AB_TW) {
        uint32_t tgt[AB_MTW][4][4];
#pragma unroll
        for (int j = 0; j < AB_MTW; ++j)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int rs = (wave + 4 * j) * 16 + 4 * (lane >> 4) + v;  // f32-output row
                    const int64_t Ib = c0 + nt * 16 + (lane & 15);
                    tgt[j][nt][v] = (rs < np && Ib < Nb) ? out_b[int64_t(rs) * Nb + Ib] : kInvalid;
                }
        for (int i = tid; i < KP * AB_TW; i += 256) {
            int e = i / AB_TW, c = i % AB_TW;
            float v = 0.f;
            if (c0 + c < Nb) {
                if (e < ns) {
                    uint32_t t = ent[e];
                    v = C[int64_t(t & kAddrMask) * Nb + c0 + c];
                    if (t >> 31) v = -v;
                } else if (e == ns) {
                    v = C[Ja * Nb + c0 + c];
                }
            }
            split_bf16(v, Xh[c * KPP + e], Xl[c * KPP + e]);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < AB_MTW; ++j) {
            const int mtile = wave + 4 * j;
            if (mtile >= nmt) break;
            f4 acc[4];
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) acc[nt] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
#pragma unroll
                for (int nt = 0; nt < 4; ++nt) {
                    const int off = (nt * 16 + (lane & 15)) * KPP + ks * 32 + 8 * (lane >> 4);
                    bf8 bh = *reinterpret_cast<const bf8*>(Xh + off);
                    bf8 bl = *reinterpret_cast<const bf8*>(Xl + off);
                    acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vl[j][ks], bh, acc[nt], 0, 0, 0);
                    acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vh[j][ks], bl, acc[nt], 0, 0, 0);
                    acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vh[j][ks], bh, acc[nt], 0, 0, 0);
                }
            }
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const uint32_t t = tgt[j][nt][v];
                    if (t == kInvalid) continue;
                    float y = (t >> 31) ? -acc[nt][v] : acc[nt][v];
                    if (use_lds_row)
                        atomicAdd(&srow[t & kAddrMask], y);
                    else
                        atomicAdd(&S[Ja * Nb + (t & kAddrMask)], y);
                }
            }
        }
        __syncthreads();
    }
    if (use_lds_row)
        for (int64_t i = tid; i < Nb; i += 256) S[Ja * Nb + i] += srow[i];
}

// ---- launchers -----------------------------------------------------------------------------
struct SigmaWork {
    DBuf<double> t64a, t64b;
    DBuf<float> t32a, t32b;
    hipEvent_t ev[4];
    bool ev_ok = false;
};

SigmaEngine::SigmaEngine(const SpinTables* a, const SpinTables* b, const DevInts* ints)
    : a_(a), b_(b), ints_(ints), w_(new SigmaWork) {
    for (int i = 0; i < 4; ++i) HIPCHK(hipEventCreate(&w_->ev[i]));
    w_->ev_ok = true;
}
SigmaEngine::~SigmaEngine() {
    if (w_ && w_->ev_ok)
        for (int i = 0; i < 4; ++i) (void)hipEventDestroy(w_->ev[i]);
}

void SigmaEngine::timings(double* ms3) const {
    for (int i = 0; i < 3; ++i) ms3[i] = ms_[i];
}

template <class T> static DBuf<T>& scratch(SigmaWork& w, int which);
template <> DBuf<double>& scratch<double>(SigmaWork& w, int which) { return which ? w.t64b : w.t64a; }
template <> DBuf<float>& scratch<float>(SigmaWork& w, int which) { return which ? w.t32b : w.t32a; }

template <class T>
static void launch_clique(const T* X, T* Y, int64_t L, const SpinTables& sp, const DevInts& ints, int spin,
                          hipStream_t st) {
    if (sp.d == 0 || sp.nK == 0) return;
    {
        // operator-in-LDS variant when it fits
        const int gd = ints.gdim[spin];
        const int Mp = (sp.M + 15) & ~15, Mk = (sp.M + 31) & ~31;
        constexpr int TW64 = 16, TW32 = 32, WV64 = 4, WV32 = 8;
        size_t sh = sizeof(T) == 8
                        ? size_t(gd) * gd * 8 + size_t(WV64) * Mp * TW64 * 8 + size_t(WV64) * Mp * 4
                        : ((size_t(gd) * gd + 3) & ~size_t(3)) * 4 + size_t(WV32) * 2 * TW32 * (Mk + 8) * 2 +
                              size_t(WV32) * Mk * 4;
        if (sh <= 150 * 1024 && sp.M <= 1023) {
            const int tw = sizeof(T) == 8 ? TW64 : TW32;
            const int nct = int((L + tw - 1) / tw);
            const int nkc = int((sp.nK + CQ2_KCH - 1) / CQ2_KCH);
            const int64_t nwork = int64_t(nct) * nkc;
            if constexpr (sizeof(T) == 8) {
                auto kern = k_clique2_f64<TW64, WV64>;
                HIPCHK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, int(sh)));
                kern<<<xcd_grid(nwork), 64 * WV64, sh, st>>>(X, Y, L, sp.clq.p, sp.M, sp.nK, ints.g2[spin].p, gd,
                                                            nct, nkc);
            } else {
                auto kern = k_clique2_mixed<TW32, WV32>;
                HIPCHK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, int(sh)));
                kern<<<xcd_grid(nwork), 64 * WV32, sh, st>>>(X, Y, L, sp.clq.p, sp.M, sp.nK, ints.g2f[spin].p, gd,
                                                            nct, nkc);
            }
            HIPCHK(hipGetLastError());
            return;
        }
    }
    if (sp.M > CQ_MAXM) throw Error(-1, "clique size exceeds kernel limit");
    int nct = int((L + CQ_TW - 1) / CQ_TW), nmt = (sp.M + CQ_MT - 1) / CQ_MT;
    int64_t nwork = int64_t(nct) * sp.nK * nmt;
    if constexpr (sizeof(T) == 8)
        k_clique_f64<<<xcd_grid(nwork), 256, 0, st>>>(X, Y, L, sp.clq.p, sp.M, sp.nK, ints.g2[spin].p,
                                                      ints.gdim[spin], nct, nmt);
    else
        k_clique_mixed<<<xcd_grid(nwork), 256, 0, st>>>(X, Y, L, sp.clq.p, sp.M, sp.nK, ints.g2f[spin].p,
                                                        ints.gdim[spin], nct, nmt);
    HIPCHK(hipGetLastError());
}

template <int KS>
static void ab_f64_launch(const double* C, double* S, const SpinTables& a, const SpinTables& b,
                          const DevInts& ints, hipStream_t st) {
    int nmt = (b.np + 15) / 16;
    size_t xs = size_t(KS * 4) * AB_TW * sizeof(double);
    size_t row = size_t(b.nstr) * sizeof(double);
    int use_row = (xs + row) <= 150 * 1024;
    size_t sh = xs + (use_row ? row : 0);
    auto kern = k_ab_f64<KS>;
    HIPCHK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, int(sh)));
    kern<<<xcd_grid(a.nstr), 256, sh, st>>>(C, S, a.nstr, b.nstr, a.into.p, a.ns, a.str.p, b.out.p, b.np,
                                            ints.vpk.p, nmt, use_row);
    HIPCHK(hipGetLastError());
}
template <int KS>
static void ab_mixed_launch(const float* C, float* S, const SpinTables& a, const SpinTables& b,
                            const DevInts& ints, hipStream_t st) {
    int nmt = (b.np + 15) / 16;
    size_t xs = size_t(AB_TW) * (KS * 32 + 8) * 2 * sizeof(__bf16);
    size_t row = size_t(b.nstr) * sizeof(float);
    int use_row = (xs + row) <= 150 * 1024;
    size_t sh = xs + (use_row ? row : 0);
    auto kern = k_ab_mixed<KS>;
    HIPCHK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, int(sh)));
    kern<<<xcd_grid(a.nstr), 256, sh, st>>>(C, S, a.nstr, b.nstr, a.into.p, a.ns, a.str.p, b.out.p, b.np,
                                            ints.vpk32.p, nmt, use_row);
    HIPCHK(hipGetLastError());
}

static void launch_ab(const double* C, double* S, const SpinTables& a, const SpinTables& b,
                      const DevInts& ints, hipStream_t st) {
    int ks = (a.ns + 1 + 3) / 4;
    if ((b.np + 15) / 16 > 4 * AB_MTW) throw Error(-1, "too many orbitals for the alpha-beta kernel");
    switch (ks) {
#define CASE(K) \
    case K: ab_f64_launch<K>(C, S, a, b, ints, st); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10)
        CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16) CASE(17) CASE(18) CASE(19) CASE(20)
        CASE(21) CASE(22) CASE(23) CASE(24) CASE(25) CASE(26) CASE(27) CASE(28) CASE(29) CASE(30)
        CASE(31) CASE(32) CASE(33) CASE(34) CASE(35) CASE(36) CASE(37) CASE(38) CASE(39) CASE(40)
#undef CASE
        default: throw Error(-1, "alpha-beta kernel: too many single excitations");
    }
}
static void launch_ab(const float* C, float* S, const SpinTables& a, const SpinTables& b,
                      const DevInts& ints, hipStream_t st) {
    int ks = (a.ns + 1 + 31) / 32;
    if ((b.np + 15) / 16 > 4 * AB_MTW) throw Error(-1, "too many orbitals for the alpha-beta kernel");
    switch (ks) {
        case 1: ab_mixed_launch<1>(C, S, a, b, ints, st); break;
        case 2: ab_mixed_launch<2>(C, S, a, b, ints, st); break;
        case 3: ab_mixed_launch<3>(C, S, a, b, ints, st); break;
        case 4: ab_mixed_launch<4>(C, S, a, b, ints, st); break;
        case 5: ab_mixed_launch<5>(C, S, a, b, ints, st); break;
        default: throw Error(-1, "alpha-beta kernel: too many single excitations");
    }
}

template <class T> void SigmaEngine::apply_t(const T* C, T* S, hipStream_t st) {
    const SpinTables& a = *a_;
    const SpinTables& b = *b_;
    const int64_t Na = a.nstr, Nb = b.nstr, nd = Na * Nb;
    const T ecore = T(ints_->e_core);
    if (timing_) HIPCHK(hipEventRecord(w_->ev[0], st));
    // beta same-spin on C^T, then sigma = (Y^T) + e_core C   (H0, fci_vector_hamiltonian.cc:84-91)
    if (b.d > 0) {
        DBuf<T>& Ct = scratch<T>(*w_, 0);
        DBuf<T>& Yt = scratch<T>(*w_, 1);
        if (Ct.n < size_t(nd)) Ct.alloc(nd);
        if (Yt.n < size_t(nd)) Yt.alloc(nd);
        transpose_device<T>(C, Ct.p, Na, Nb, st);
        k_fill<T><<<nblk(nd, 256), 256, 0, st>>>(Yt.p, nd, T(0));
        launch_clique<T>(Ct.p, Yt.p, Na, b, *ints_, 1, st);
        transpose_add_device<T>(Yt.p, S, Nb, Na, C, ecore, st);
    } else {
        k_scale_copy<T><<<nblk(nd, 256), 256, 0, st>>>(C, S, nd, ecore);
    }
    if (timing_) HIPCHK(hipEventRecord(w_->ev[1], st));
    if (a.d > 0) launch_clique<T>(C, S, Nb, a, *ints_, 0, st);
    if (timing_) HIPCHK(hipEventRecord(w_->ev[2], st));
    if (a.k > 0 && b.k > 0) launch_ab(C, S, a, b, *ints_, st);
    if (timing_) {
        HIPCHK(hipEventRecord(w_->ev[3], st));
        HIPCHK(hipEventSynchronize(w_->ev[3]));
        for (int i = 0; i < 3; ++i) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, w_->ev[i], w_->ev[i + 1]));
            ms_[i] = ms;
        }
    }
}

void SigmaEngine::apply(const double* C, double* S, hipStream_t st) { apply_t<double>(C, S, st); }
void SigmaEngine::apply(const float* C, float* S, hipStream_t st) { apply_t<float>(C, S, st); }

}  // namespace fci
