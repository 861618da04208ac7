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
#include <cstdlib>
#include <cstring>
#include <map>

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

// 64x64 tiles with two-element loads and stores (even row and column counts): every load and
// store instruction of a wave covers two 64-element row segments.
template <class T> struct Pair;
template <> struct Pair<float> { using type = float2; };
template <> struct Pair<double> { using type = double2; };
// Batched launches (many problems of one size at once) pass per-item pointer tables indexed by
// blockIdx.y; an item whose `act` entry is null is idle. Null tables: the plain pointers are used.
template <class T>
__global__ __launch_bounds__(256) void k_transpose64(const T* __restrict__ in, T* __restrict__ out, int64_t R,
                                                     int64_t Cc, const T* __restrict__ base, T bs,
                                                     const void* const* __restrict__ act = nullptr,
                                                     const T* const* __restrict__ in_t = nullptr,
                                                     T* const* __restrict__ out_t = nullptr,
                                                     const T* const* __restrict__ base_t = nullptr) {
    using V2 = typename Pair<T>::type;
    __shared__ T tile[64][65];
    if (act) {
        if (!act[blockIdx.z]) return;
        in = in_t[blockIdx.z];
        out = out_t[blockIdx.z];
        if (base_t) base = base_t[blockIdx.z];
    }
    const int64_t c0 = int64_t(blockIdx.x) * 64, r0 = int64_t(blockIdx.y) * 64;
    const int cp = threadIdx.x & 31, rr = threadIdx.x >> 5;  // column pair, row in a group of 8
    for (int j = rr; j < 64; j += 8) {
        const int64_t r = r0 + j, c = c0 + 2 * cp;  // c even and Cc even: c + 1 < Cc as well
        if (r < R && c < Cc) {
            const V2 v = *reinterpret_cast<const V2*>(in + r * Cc + c);
            tile[j][2 * cp] = v.x;
            tile[j][2 * cp + 1] = v.y;
        }
    }
    __syncthreads();
    for (int j = rr; j < 64; j += 8) {
        const int64_t c = c0 + j, r = r0 + 2 * cp;  // output row c, columns r and r + 1
        if (c < Cc && r < R) {
            V2 v;
            v.x = tile[2 * cp][j];
            v.y = tile[2 * cp + 1][j];
            if (base) {
                const V2 b = *reinterpret_cast<const V2*>(base + c * R + r);
                v.x += bs * b.x;
                v.y += bs * b.y;
            }
            *reinterpret_cast<V2*>(out + c * R + r) = v;
        }
    }
}

template <class T>
static void transpose_launch(const T* in, T* out, int64_t rows, int64_t cols, const T* base, T bs, hipStream_t st) {
    if (rows % 2 == 0 && cols % 2 == 0) {
        dim3 g(nblk(cols, 64), nblk(rows, 64));
        k_transpose64<T><<<g, 256, 0, st>>>(in, out, rows, cols, base, bs);
    } else {
        dim3 g(nblk(cols, 32), nblk(rows, 32)), b(32, 8);
        k_transpose<T><<<g, b, 0, st>>>(in, out, rows, cols, base, bs);
    }
    HIPCHK(hipGetLastError());
}
template <class T> void transpose_device(const T* in, T* out, int64_t rows, int64_t cols, hipStream_t st) {
    transpose_launch<T>(in, out, rows, cols, nullptr, T(0), st);
}
template <class T>
void transpose_add_device(const T* in, T* out, int64_t rows, int64_t cols, const T* base, T bs,
                          hipStream_t st) {
    transpose_launch<T>(in, out, rows, cols, base, bs, st);
}
template void transpose_device<double>(const double*, double*, int64_t, int64_t, hipStream_t);
template void transpose_device<float>(const float*, float*, int64_t, int64_t, hipStream_t);
template void transpose_add_device<double>(const double*, double*, int64_t, int64_t, const double*, double, hipStream_t);
template void transpose_add_device<float>(const float*, float*, int64_t, int64_t, const float*, float, hipStream_t);

// S <- S + eta S^T in place (optionally halved), tile pair (I, J), I <= J, per block: both 64 x 64
// tiles are read into LDS before either is written. Element (i, j) becomes a + eta b and its mirror
// b + eta a with a = S[i][j], b = S[j][i]; IEEE addition is commutative and x - y = -(y - x), so
// the mirror is exactly eta times the element.
template <class T>
__global__ __launch_bounds__(256) void k_sym_add(T* __restrict__ S, int64_t N, int eta, int avg) {
    const int I = blockIdx.y, J = blockIdx.x;
    if (J < I) return;
    __shared__ T ta[64][65], tb[64][65];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int64_t r0 = int64_t(I) * 64, c0 = int64_t(J) * 64;
    const T e = T(eta), h = avg ? T(0.5) : T(1);
    for (int j = ty; j < 64; j += 4) {
        const int64_t r = r0 + j, c = c0 + tx;
        if (r < N && c < N) ta[j][tx] = S[r * N + c];  // tile (I, J)
        const int64_t r2 = c0 + j, c2 = r0 + tx;
        if (I != J && r2 < N && c2 < N) tb[j][tx] = S[r2 * N + c2];  // tile (J, I)
    }
    __syncthreads();
    for (int j = ty; j < 64; j += 4) {
        const int64_t r = r0 + j, c = c0 + tx;
        if (r < N && c < N) {
            const T a = ta[j][tx], b = I != J ? tb[tx][j] : ta[tx][j];
            S[r * N + c] = h * (a + e * b);
        }
        const int64_t r2 = c0 + j, c2 = r0 + tx;
        if (I != J && r2 < N && c2 < N) {
            const T a = tb[j][tx], b = ta[tx][j];
            S[r2 * N + c2] = h * (a + e * b);
        }
    }
}

template <class T> void sym_add_device(T* S, int64_t N, int eta, bool avg, hipStream_t st) {
    if (N <= 0) return;
    const unsigned t = unsigned((N + 63) / 64);
    k_sym_add<T><<<dim3(t, t), 256, 0, st>>>(S, N, eta, avg ? 1 : 0);
    HIPCHK(hipGetLastError());
}
template void sym_add_device<double>(double*, int64_t, int, bool, hipStream_t);
template void sym_add_device<float>(float*, int64_t, int, bool, hipStream_t);

// ---- spin-flip-symmetric vectors packed to the upper triangle ------------------------------------
// Row i of the packed layout holds columns j = i + d .. N - 1 (d = 0: diagonal included, d = 1:
// excluded), at offset i (N - d) - i (i - 1) / 2.
__device__ __host__ inline int64_t ms0_row_off(int64_t i, int64_t N, int d) { return i * (N - d) - i * (i - 1) / 2; }
int64_t ms0_packed_size(int64_t N, int eta) { return eta == 1 ? N * (N + 1) / 2 : N * (N - 1) / 2; }
int64_t ms0_packed_index(int64_t i, int64_t j, int64_t N, int eta) {
    const int d = eta == 1 ? 0 : 1;
    return ms0_row_off(i, N, d) + (j - i - d);
}
std::vector<std::pair<int64_t, double>> ms0_pack_sparse(const std::vector<std::pair<int64_t, double>>& v,
                                                        int64_t N, int eta) {
    std::map<int64_t, double> m;
    for (const auto& [i, c] : v) m[i] = c;  // later duplicates overwrite, as set_vector
    std::map<int64_t, double> out;
    for (const auto& [k, c] : m) {
        const int64_t i = k / N, j = k % N;
        if (i == j) {
            if (eta == 1) out[ms0_packed_index(i, i, N, eta)] = c;
            continue;
        }
        // C_ij enters the pair (min, max) directly (i < j) or through its mirror's sign (i > j)
        const int64_t lo = std::min(i, j), hi = std::max(i, j);
        out[ms0_packed_index(lo, hi, N, eta)] += (i < j ? c : eta * c) * 0.70710678118654752440;
    }
    return std::vector<std::pair<int64_t, double>>(out.begin(), out.end());
}

// packed(i, j) = s_off (F_ij + mirror F_ji) for i < j, s_diag F_ii on the diagonal (d = 0); tile
// pair (I, J), I <= J, per block, both tiles of F read into LDS
// Batched launches (many problems of one size) pass per-item pointer tables indexed by blockIdx.z;
// an item whose `act` entry is null is idle. Null tables: the plain pointers are used.
template <class TI, class TO>
__global__ __launch_bounds__(256) void k_ms0_pack(const TI* __restrict__ F, TO* __restrict__ P, int64_t N, int d,
                                                  int mirror, double s_off, double s_diag,
                                                  const void* const* __restrict__ act = nullptr,
                                                  const TI* const* __restrict__ F_t = nullptr,
                                                  TO* const* __restrict__ P_t = nullptr) {
    const int I = blockIdx.y, J = blockIdx.x;
    if (J < I) return;
    if (act) {
        if (!act[blockIdx.z]) return;
        F = F_t[blockIdx.z];
        P = P_t[blockIdx.z];
    }
    __shared__ TI ta[64][65], tb[64][65];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int64_t r0 = int64_t(I) * 64, c0 = int64_t(J) * 64;
    for (int j = ty; j < 64; j += 4) {
        const int64_t r = r0 + j, c = c0 + tx;
        if (r < N && c < N) ta[j][tx] = F[r * N + c];
        const int64_t r2 = c0 + j, c2 = r0 + tx;
        if (I != J && r2 < N && c2 < N) tb[j][tx] = F[r2 * N + c2];
    }
    __syncthreads();
    const TI m = TI(mirror), so = TI(s_off), sd = TI(s_diag);
    for (int j = ty; j < 64; j += 4) {
        const int64_t i = r0 + j, c = c0 + tx;
        if (i >= N || c >= N || c < i + d) continue;
        const TI a = ta[j][tx];
        TI v;
        if (c == i) {
            v = sd * a;
        } else {
            const TI b = I != J ? tb[tx][j] : ta[tx][j];
            v = so * (a + m * b);
        }
        P[ms0_row_off(i, N, d) + (c - i - d)] = TO(v);
    }
}

// F_ij = P(i, j) / sqrt2 and F_ji = eta F_ij for i < j, F_ii = P(i, i) (d = 0) or 0 (d = 1); a
// mirrored pair is converted once, so F is exactly eta-symmetric in its own precision
template <class TI, class TO>
__global__ __launch_bounds__(256) void k_ms0_unpack(const TI* __restrict__ P, TO* __restrict__ F, int64_t N, int d,
                                                    int eta, TO* __restrict__ S, TO s_scale,
                                                    const void* const* __restrict__ act = nullptr,
                                                    const TI* const* __restrict__ P_t = nullptr,
                                                    TO* const* __restrict__ F_t = nullptr) {
    const int I = blockIdx.y, J = blockIdx.x;
    if (J < I) return;
    if (act) {
        if (!act[blockIdx.z]) return;
        P = P_t[blockIdx.z];
        F = F_t[blockIdx.z];
    }
    __shared__ TO tp[64][65];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int64_t r0 = int64_t(I) * 64, c0 = int64_t(J) * 64;
    const TI h = TI(0.70710678118654752440);
    for (int j = ty; j < 64; j += 4) {
        const int64_t i = r0 + j, c = c0 + tx;
        TO v = TO(0);
        if (i < N && c < N && c >= i + d) {
            const TI x = P[ms0_row_off(i, N, d) + (c - i - d)];
            v = c == i ? TO(x) : TO(x * h);
        }
        tp[j][tx] = v;
    }
    __syncthreads();
    const TO e = TO(eta);
    for (int j = ty; j < 64; j += 4) {
        const int64_t i = r0 + j, c = c0 + tx;
        if (i < N && c < N) {
            // upper part of tile (I, J); in the diagonal tile the lower part mirrors the upper one
            TO v;
            if (c >= i) v = tp[j][tx];
            else v = e * tp[tx][j];
            F[i * N + c] = v;
            if (S) S[i * N + c] = s_scale * v;
        }
        const int64_t i2 = c0 + j, c2 = r0 + tx;  // tile (J, I): the mirror
        if (I != J && i2 < N && c2 < N) {
            const TO v = e * tp[tx][j];
            F[i2 * N + c2] = v;
            if (S) S[i2 * N + c2] = s_scale * v;
        }
    }
}

__global__ void k_ms0_pack_mask(const uint8_t* __restrict__ F, uint8_t* __restrict__ P, int64_t N, int d) {
    const int64_t c = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (c >= N) return;
    for (int64_t i = blockIdx.y; i < N; i += gridDim.y)  // rows strided over the grid's y extent
        if (c >= i + d) P[ms0_row_off(i, N, d) + (c - i - d)] = F[i * N + c];
}

template <class TI, class TO>
void ms0_pack_device(const TI* full, TO* packed, int64_t N, int eta, int mirror, double s_off, double s_diag,
                     hipStream_t st) {
    const unsigned t = unsigned((N + 63) / 64);
    k_ms0_pack<TI, TO><<<dim3(t, t), 256, 0, st>>>(full, packed, N, eta == 1 ? 0 : 1, mirror, s_off, s_diag);
    HIPCHK(hipGetLastError());
}
template <class TI, class TO>
void ms0_unpack_device(const TI* packed, TO* full, int64_t N, int eta, hipStream_t st, TO* s_init, double s_scale) {
    const unsigned t = unsigned((N + 63) / 64);
    k_ms0_unpack<TI, TO><<<dim3(t, t), 256, 0, st>>>(packed, full, N, eta == 1 ? 0 : 1, eta, s_init, TO(s_scale));
    HIPCHK(hipGetLastError());
}
void ms0_pack_mask_device(const uint8_t* full, uint8_t* packed, int64_t N, int eta, hipStream_t st) {
    const unsigned gy = unsigned(std::min<int64_t>(N, 65535));
    k_ms0_pack_mask<<<dim3(unsigned((N + 255) / 256), gy), 256, 0, st>>>(full, packed, N, eta == 1 ? 0 : 1);
    HIPCHK(hipGetLastError());
}
template void ms0_pack_device<double, double>(const double*, double*, int64_t, int, int, double, double, hipStream_t);
template void ms0_pack_device<float, float>(const float*, float*, int64_t, int, int, double, double, hipStream_t);
template void ms0_unpack_device<double, double>(const double*, double*, int64_t, int, hipStream_t, double*, double);
template void ms0_unpack_device<double, float>(const double*, float*, int64_t, int, hipStream_t, float*, double);

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

// Column-tile-major variant: each wave owns one hole string K at a time (kch hole strings per
// block) and the XCD-contiguous work order keeps one column tile of X resident in that XCD's L2
// while every K of the tile passes through it. The MFMA A operand is read from g' through the
// member pair indices, so no per-K copy of W_K is made (FP64: g' through L1; mixed: g' in LDS).

// STORE = true: no atomics; E[(K*M + m) * Ls + (col - cbeg)] = s_m sum_m' W s_m' X (pass 1 of the
// two-pass same-spin product, gathered by k_clique_gather).
template <int TW, int CQ2_WAVES, bool STORE>
__global__ __launch_bounds__(64 * CQ2_WAVES) void k_clique2_f64(const double* __restrict__ X, double* __restrict__ Y,
                                                                int64_t L, const uint32_t* __restrict__ clq, int M,
                                                                int64_t nK, const double* __restrict__ G, int gdim,
                                                                int nct, int nkc, int kch, int64_t cbeg, int64_t Ls) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    constexpr int NT = TW / 16;
    const double* Gs = G;                                             // read through L1/L2
    const int Mp = (M + 15) & ~15;
    double* Dall = reinterpret_cast<double*>(smem);                   // per wave [Mp][TW]
    uint32_t* memall = reinterpret_cast<uint32_t*>(Dall + size_t(CQ2_WAVES) * Mp * TW);  // per wave [Mp]
    const int64_t nwork = int64_t(nct) * nkc;
    const int64_t w = xcd_remap(blockIdx.x, nwork);
    if (w >= nwork) return;
    const int kc = int(w % nkc), ct = int(w / nkc);
    const int64_t c0 = cbeg + int64_t(ct) * TW;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double* Ds = Dall + size_t(wave) * Mp * TW;
    uint32_t* mem = memall + size_t(wave) * Mp;
    const int nmt = Mp / 16;
    for (int64_t K = int64_t(kc) * kch + wave; K < min(int64_t(kc + 1) * kch, nK); K += CQ2_WAVES) {
        for (int i = lane; i < Mp; i += 64) mem[i] = i < M ? clq[K * M + i] : 0u;
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        for (int i = lane; i < Mp * TW; i += 64) {
            int m = i / TW, c = i % TW;
            double v = 0.0;
            if (m < M && c0 + c < L && c0 + c < cbeg + Ls) {
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
                    if (r < M && col < cbeg + Ls && col < L) {
                        uint32_t e = mem[r];
                        const double y = (e >> 31) ? -acc[nt][v] : acc[nt][v];
                        if constexpr (STORE)
                            Y[(K * M + r) * Ls + (col - cbeg)] = y;
                        else
                            atomicAdd(&Y[int64_t(e & kAddrMask) * L + col], y);
                    }
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

template <int TW, int CQ2_WAVES, bool STORE>
__global__ __launch_bounds__(64 * CQ2_WAVES) void k_clique2_mixed(const float* __restrict__ X, float* __restrict__ Y,
                                                                  int64_t L, const uint32_t* __restrict__ clq, int M,
                                                                  int64_t nK, const float* __restrict__ G, int gdim,
                                                                  int nct, int nkc, int kch, int64_t cbeg, int64_t Ls) {
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
    const int64_t c0 = cbeg + int64_t(ct) * TW;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < gdim * gdim; i += 64 * CQ2_WAVES) Gs[i] = G[i];
    __syncthreads();
    __bf16* Dh = Dall + size_t(wave) * 2 * TW * KPP;
    __bf16* Dl = Dh + TW * KPP;
    uint32_t* mem = memall + size_t(wave) * Mk;
    const int nmt = Mp / 16;
    for (int64_t K = int64_t(kc) * kch + wave; K < min(int64_t(kc + 1) * kch, nK); K += CQ2_WAVES) {
        for (int i = lane; i < Mk; i += 64) mem[i] = i < M ? clq[K * M + i] : 0u;
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        for (int i = lane; i < Mk * TW; i += 64) {
            int m = i / TW, c = i % TW;
            float v = 0.f;
            if (m < M && c0 + c < L && c0 + c < cbeg + Ls) {
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
                    if (r < M && col < cbeg + Ls && col < L) {
                        uint32_t e = mem[r];
                        const float y = (e >> 31) ? -acc[nt][v] : acc[nt][v];
                        if constexpr (STORE)
                            Y[(K * M + r) * Ls + (col - cbeg)] = y;
                        else
                            atomicAdd(&Y[int64_t(e & kAddrMask) * L + col], y);
                    }
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// Pass 1 of the two-pass same-spin product for cliques of at most 64 members, with no LDS: each
// wave owns one (hole string K, 64-column chunk) work item, keeps the MFMA A operand
// W_K[m][m'] = g'[pair(m)][pair(m')] in registers and streams 16-column tiles of
// X[I_m'][c] straight from global memory into the B operand layout, one tile ahead.
// E[(K * M + m) * Ls + c - cbeg] = s_m sum_m' W_K[m][m'] s_m' X[I_m'][c].
// Row of a gathered string for an unconditional load: an invalid entry reads row 0, whose
// contribution the zero A-operand column of that entry cancels. Keeping the gather loads free of
// per-lane branches lets all of a tile's loads be in flight at once (a load whose result is used
// inside a divergent branch is waited for on the spot).
__device__ __forceinline__ int64_t gather_row(uint32_t e) { return e != kInvalid ? int64_t(e & kAddrMask) : 0; }

constexpr int CQ3_CW = 64;  // columns per work item

template <int NMT, bool GL>  // row tiles of 16 members; k steps of 32 members: (NMT + 1) / 2; GL: g' in LDS
__global__ __launch_bounds__(256) void k_clique3_mixed(const float* __restrict__ X, float* __restrict__ E, int64_t L,
                                                       const uint32_t* __restrict__ clq, int M, int64_t nK,
                                                       const float* __restrict__ Gg, int gdim, int nchunk,
                                                       int64_t cbeg, int64_t Ls) {
    constexpr int NKS = (NMT + 1) / 2;
    extern __shared__ __attribute__((aligned(16))) float gl_f32[];
    const float* G = Gg;
    if constexpr (GL) {
        for (int i = threadIdx.x; i < gdim * gdim; i += 256) gl_f32[i] = Gg[i];
        __syncthreads();
        G = gl_f32;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, cl = lane & 15;
    const int64_t nwork = nK * nchunk;
    const int64_t nblk = (nwork + 3) / 4;
    const int64_t w = xcd_remap(blockIdx.x, nblk) * 4 + wave;
    if (w >= nwork) return;
    const int64_t K = w % nK;
    const int chunk = int(w / nK);
    const uint32_t* ent = clq + K * M;
    uint32_t mb[NKS][8];
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks)
#pragma unroll
        for (int x = 0; x < 8; ++x) {
            const int m = ks * 32 + 8 * g + x;
            const uint32_t t = ent[m < M ? m : 0];  // unconditional loads (M >= 1): all in flight at once
            mb[ks][x] = m < M ? t : kInvalid;
        }
    bf8 ah[NMT][NKS], al[NMT][NKS];
    uint32_t neg[NMT];  // sign bits of this lane's 4 output rows, bit v
#pragma unroll
    for (int mt = 0; mt < NMT; ++mt) {
        const int r = mt * 16 + cl;
        const int pr = int((ent[r < M ? r : 0] >> 21) & 0x3FFu);  // rows r >= M get zero below
#pragma unroll
        for (int ks = 0; ks < NKS; ++ks)
#pragma unroll
            for (int x = 0; x < 8; ++x) {
                const uint32_t e = mb[ks][x];
                const float t = G[pr * gdim + (e != kInvalid ? int((e >> 21) & 0x3FFu) : 0)];
                float v = (r < M && e != kInvalid) ? t : 0.f;
                if (e >> 31) v = -v;  // sign of the gathered row folded into A
                __bf16 h, l;
                split_bf16(v, h, l);
                ah[mt][ks][x] = h;
                al[mt][ks][x] = l;
            }
        uint32_t nb = 0;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int ro = mt * 16 + 4 * g + v;
            const uint32_t t = ent[ro < M ? ro : 0];
            nb |= (ro < M ? t >> 31 : 0u) << v;
        }
        neg[mt] = nb;
    }
    const int64_t cend = cbeg + Ls < L ? cbeg + Ls : L;
    const int64_t c0 = cbeg + int64_t(chunk) * CQ3_CW;
    float xv[NKS][8], xn[NKS][8];
    auto load = [&](int64_t col, float (&xx)[NKS][8]) {
        const int64_t cc = col < cend ? col : cbeg;  // columns past the end are computed, not stored
#pragma unroll
        for (int ks = 0; ks < NKS; ++ks)
#pragma unroll
            for (int x = 0; x < 8; ++x) xx[ks][x] = X[gather_row(mb[ks][x]) * L + cc];
    };
    load(c0 + cl, xv);
#pragma unroll 1
    for (int ct = 0; ct < CQ3_CW / 16; ++ct) {
        const int64_t col = c0 + ct * 16 + cl;
        if (ct + 1 < CQ3_CW / 16) load(col + 16, xn);
        bf8 bh[NKS], bl[NKS];
#pragma unroll
        for (int ks = 0; ks < NKS; ++ks)
#pragma unroll
            for (int x = 0; x < 8; ++x) {
                __bf16 h, l;
                split_bf16(xv[ks][x], h, l);
                bh[ks][x] = h;
                bl[ks][x] = l;
            }
#pragma unroll
        for (int mt = 0; mt < NMT; ++mt) {
            f4 acc = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < NKS; ++ks) {
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[mt][ks], bh[ks], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[mt][ks], bl[ks], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[mt][ks], bh[ks], acc, 0, 0, 0);
            }
            if (col < cend) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = mt * 16 + 4 * g + v;
                    if (r < M)
                        E[((int64_t(chunk) * nK + K) * M + r) * CQ3_CW + (col - c0)] =
                            ((neg[mt] >> v) & 1u) ? -acc[v] : acc[v];
                }
            }
        }
#pragma unroll
        for (int ks = 0; ks < NKS; ++ks)
#pragma unroll
            for (int x = 0; x < 8; ++x) xv[ks][x] = xn[ks][x];
    }
}

// Paired-column variant for even row lengths: each lane loads two adjacent columns with one 8-byte
// load, which feeds two MFMA column tiles (columns c0 + 2j and c0 + 2j + 1, j = lane % 16), and
// stores both results with one 8-byte store: half the memory instructions of k_clique3_mixed.
// MS > 0: the clique size M = MS is known at compile time. Member m sits in k slot 8g + x of
// k-step ks with m = 32 ks + 4 x + g, so the padding m >= M fills whole x rows of the last k-step,
// whose gathers (and their member-table loads) are then left out.
template <int NMT, int MS>
__global__ __launch_bounds__(256) void k_clique3p_mixed(const float* __restrict__ X, float* __restrict__ E, int64_t L,
                                                        const uint32_t* __restrict__ clq, int Mrt, int64_t nK,
                                                        const float* __restrict__ G, int gdim, int nchunk,
                                                        int64_t cbeg, int64_t Ls, int cpw) {
    constexpr int NKS = (NMT + 1) / 2;
    const int M = MS > 0 ? MS : Mrt;
    // k slots (ks, x) with no member for any g (only when M is static)
    auto absent = [](int ks, int x) { return MS > 0 && ks * 32 + 4 * x >= MS; };
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, cl = lane & 15;
    // MS > 0: a chunk's E rows are staged in this wave's LDS rows (stride CQ_ES floats) and
    // written to E as whole 256-byte rows with 16-byte stores, half the store instructions
    constexpr int CQ_ES = 68;
    __shared__ __attribute__((aligned(16))) float et_all[MS > 0 ? 4 : 1][MS > 0 ? NMT * 16 * CQ_ES : 4];
    float* et = et_all[MS > 0 ? wave : 0];
    const int ngrp = (nchunk + cpw - 1) / cpw;  // work item: one K and up to cpw consecutive chunks
    const int64_t nwork = nK * ngrp;
    const int64_t nblk = (nwork + 3) / 4;
    const int64_t w = xcd_remap(blockIdx.x, nblk) * 4 + wave;
    if (w >= nwork) return;
    const int64_t K = w % nK;
    const int ch0 = int(w / nK) * cpw;
    const int nh = 2 * min(cpw, nchunk - ch0);  // 32-column halves of this work item
    const uint32_t* ent = clq + K * M;
    uint32_t mb[NKS][8];
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks)
#pragma unroll
        for (int x = 0; x < 8; ++x) {
            const int m = ks * 32 + 4 * x + g;
            if (absent(ks, x)) {
                mb[ks][x] = kInvalid;
                continue;
            }
            const uint32_t t = ent[m < M ? m : 0];  // unconditional loads (M >= 1): all in flight at once
            mb[ks][x] = m < M ? t : kInvalid;
        }
    bf8 ah[NMT][NKS], al[NMT][NKS];
    uint32_t neg[NMT];
#pragma unroll
    for (int mt = 0; mt < NMT; ++mt) {
        const int r = mt * 16 + cl;
        const int pr = int((ent[r < M ? r : 0] >> 21) & 0x3FFu);  // rows r >= M get zero below
#pragma unroll
        for (int ks = 0; ks < NKS; ++ks)
#pragma unroll
            for (int x = 0; x < 8; ++x) {
                const uint32_t e = mb[ks][x];
                const float t = G[pr * gdim + (e != kInvalid ? int((e >> 21) & 0x3FFu) : 0)];
                float v = (r < M && e != kInvalid) ? t : 0.f;
                if (e >> 31) v = -v;  // sign of the gathered row folded into A
                __bf16 h, l;
                split_bf16(v, h, l);
                ah[mt][ks][x] = h;
                al[mt][ks][x] = l;
            }
        uint32_t nb = 0;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int ro = mt * 16 + 4 * g + v;
            const uint32_t t = ent[ro < M ? ro : 0];
            nb |= (ro < M ? t >> 31 : 0u) << v;
        }
        neg[mt] = nb;
    }
    const int64_t cend = cbeg + Ls < L ? cbeg + Ls : L;  // even
    float2 xv[NKS][8], xn[NKS][8];
    auto load = [&](int64_t col, float2 (&xx)[NKS][8]) {
        const int64_t cc = col < cend ? col : cbeg;  // col, cbeg even; past the end: not stored
#pragma unroll
        for (int ks = 0; ks < NKS; ++ks)
#pragma unroll
            for (int x = 0; x < 8; ++x)
                xx[ks][x] = absent(ks, x) ? make_float2(0.f, 0.f)
                                          : *reinterpret_cast<const float2*>(X + gather_row(mb[ks][x]) * L + cc);
    };
    auto colof = [&](int h) { return cbeg + int64_t(ch0 + (h >> 1)) * CQ3_CW + (h & 1) * 32 + 2 * cl; };
    auto compute = [&](const float2 (&x)[NKS][8], int h) {
        const int chunk = ch0 + (h >> 1);
        const int64_t col = colof(h), c0 = cbeg + int64_t(chunk) * CQ3_CW;
        bf8 bh0[NKS], bl0[NKS], bh1[NKS], bl1[NKS];
#pragma unroll
        for (int ks = 0; ks < NKS; ++ks)
#pragma unroll
            for (int xx = 0; xx < 8; ++xx) {
                __bf16 h, l;
                split_bf16(x[ks][xx].x, h, l);
                bh0[ks][xx] = h;
                bl0[ks][xx] = l;
                split_bf16(x[ks][xx].y, h, l);
                bh1[ks][xx] = h;
                bl1[ks][xx] = l;
            }
#pragma unroll
        for (int mt = 0; mt < NMT; ++mt) {
            f4 a0 = f4{0.f, 0.f, 0.f, 0.f}, a1 = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < NKS; ++ks) {
                a0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[mt][ks], bh0[ks], a0, 0, 0, 0);
                a1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[mt][ks], bh1[ks], a1, 0, 0, 0);
                a0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[mt][ks], bl0[ks], a0, 0, 0, 0);
                a1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[mt][ks], bl1[ks], a1, 0, 0, 0);
                a0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[mt][ks], bh0[ks], a0, 0, 0, 0);
                a1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[mt][ks], bh1[ks], a1, 0, 0, 0);
            }
            if constexpr (MS > 0) {  // rows r >= M land in the unused staging rows
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = mt * 16 + 4 * g + v;
                    const bool ng = (neg[mt] >> v) & 1u;
                    *reinterpret_cast<float2*>(et + r * CQ_ES + (col - c0)) =
                        ng ? make_float2(-a0[v], -a1[v]) : make_float2(a0[v], a1[v]);
                }
            } else if (col < cend) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = mt * 16 + 4 * g + v;
                    if (r < M) {
                        const bool ng = (neg[mt] >> v) & 1u;
                        *reinterpret_cast<float2*>(E + ((int64_t(chunk) * nK + K) * M + r) * CQ3_CW + (col - c0)) =
                            ng ? make_float2(-a0[v], -a1[v]) : make_float2(a0[v], a1[v]);
                    }
                }
            }
        }
    };
    // one chunk from the staging rows to E. The chunk's block of E is M contiguous rows of 64
    // floats; columns past the slab end hold values of clamped loads that pass 2 never reads.
    // The LDS accesses of one wave execute in order; the compiler barriers keep the writes of the
    // two halves before these reads and these reads before the next chunk's writes.
    auto flush = [&](int chunk) {
        if constexpr (MS > 0) {
            asm volatile("" ::: "memory");
            __builtin_amdgcn_wave_barrier();
            float* eb = E + (int64_t(chunk) * nK + K) * MS * CQ3_CW;
#pragma unroll
            for (int i0 = 0; i0 < MS * 16; i0 += 64) {
                const int i = i0 + lane;
                if (i < MS * 16) {
                    const int r = i >> 4, c4 = (i & 15) * 4;
                    *reinterpret_cast<float4*>(eb + r * CQ3_CW + c4) = *reinterpret_cast<const float4*>(et + r * CQ_ES + c4);
                }
            }
            asm volatile("" ::: "memory");
            __builtin_amdgcn_wave_barrier();
        }
    };
    // halves in two buffers used alternately: the loads of the next half are in flight while the
    // current one is multiplied and stored
    load(colof(0), xv);
#pragma unroll 1
    for (int h = 0; h < nh; h += 2) {
        load(colof(h + 1), xn);
        compute(xv, h);
        if (h + 2 < nh) load(colof(h + 2), xv);
        compute(xn, h + 1);
        flush(ch0 + (h >> 1));
    }
}

template <int NK4, bool GL>  // k steps of 4 members; row tiles of 16: (NK4 + 3) / 4; GL: g' in LDS
__global__ __launch_bounds__(512) void k_clique3_f64(const double* __restrict__ X, double* __restrict__ E, int64_t L,
                                                     const uint32_t* __restrict__ clq, int M, int64_t nK,
                                                     const double* __restrict__ Gg, int gdim, int nchunk,
                                                     int64_t cbeg, int64_t Ls) {
    constexpr int NMT = (NK4 + 3) / 4;
    extern __shared__ __attribute__((aligned(16))) double gl_f64[];
    const double* G = Gg;
    if constexpr (GL) {
        for (int i = threadIdx.x; i < gdim * gdim; i += 512) gl_f64[i] = Gg[i];
        __syncthreads();
        G = gl_f64;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, cl = lane & 15;
    const int64_t nwork = nK * nchunk;
    const int64_t nblk = (nwork + 7) / 8;
    const int64_t w = xcd_remap(blockIdx.x, nblk) * 8 + wave;
    if (w >= nwork) return;
    const int64_t K = w % nK;
    const int chunk = int(w / nK);
    const uint32_t* ent = clq + K * M;
    uint32_t mb[NK4];
#pragma unroll
    for (int ks = 0; ks < NK4; ++ks) {
        const int m = ks * 4 + g;
        const uint32_t t = ent[m < M ? m : 0];  // unconditional loads (M >= 1): all in flight at once
        mb[ks] = m < M ? t : kInvalid;
    }
    double va[NMT][NK4];
    uint32_t neg[NMT];
#pragma unroll
    for (int mt = 0; mt < NMT; ++mt) {
        const int r = mt * 16 + cl;
        const int pr = int((ent[r < M ? r : 0] >> 21) & 0x3FFu);  // rows r >= M get zero below
#pragma unroll
        for (int ks = 0; ks < NK4; ++ks) {
            const uint32_t e = mb[ks];
            const double t = G[pr * gdim + (e != kInvalid ? int((e >> 21) & 0x3FFu) : 0)];
            const double v = (r < M && e != kInvalid) ? t : 0.0;
            va[mt][ks] = (e >> 31) ? -v : v;  // sign of the gathered row folded into A
        }
        uint32_t nb = 0;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int ro = mt * 16 + 4 * v + g;  // f64 output row
            const uint32_t t = ent[ro < M ? ro : 0];
            nb |= (ro < M ? t >> 31 : 0u) << v;
        }
        neg[mt] = nb;
    }
    const int64_t cend = cbeg + Ls < L ? cbeg + Ls : L;
    const int64_t c0 = cbeg + int64_t(chunk) * CQ3_CW;
    double xv[NK4], xn[NK4];
    auto load = [&](int64_t col, double (&xx)[NK4]) {
        const int64_t cc = col < cend ? col : cbeg;  // columns past the end are computed, not stored
#pragma unroll
        for (int ks = 0; ks < NK4; ++ks) xx[ks] = X[gather_row(mb[ks]) * L + cc];
    };
    auto compute = [&](const double (&xx)[NK4], int64_t col) {
#pragma unroll
        for (int mt = 0; mt < NMT; ++mt) {
            d4 acc = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int ks = 0; ks < NK4; ++ks) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(va[mt][ks], xx[ks], acc, 0, 0, 0);
            if (col < cend) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = mt * 16 + 4 * v + g;
                    if (r < M)
                        E[((int64_t(chunk) * nK + K) * M + r) * CQ3_CW + (col - c0)] =
                            ((neg[mt] >> v) & 1u) ? -acc[v] : acc[v];
                }
            }
        }
    };
    // four 16-column tiles, two buffers used alternately (no register copies)
    static_assert(CQ3_CW == 64, "four tiles per work item");
    load(c0 + cl, xv);
    load(c0 + 16 + cl, xn);
    compute(xv, c0 + cl);
    load(c0 + 32 + cl, xv);
    compute(xn, c0 + 16 + cl);
    load(c0 + 48 + cl, xn);
    compute(xv, c0 + 32 + cl);
    compute(xn, c0 + 48 + cl);
}

// FP64 pass 1 for even row lengths with paired columns: lane (g, c) loads columns 2c and 2c + 1
// of its gathered row with one 16-byte load (four rows x 256 bytes per instruction), which feed
// two MFMA column tiles, and stores both results with one 16-byte store: a 64-column chunk is two
// 32-column halves, half the memory instructions of k_clique3_f64.
template <int NK4>
__global__ __launch_bounds__(512) void k_clique3p_f64(const double* __restrict__ X, double* __restrict__ E, int64_t L,
                                                      const uint32_t* __restrict__ clq, int M, int64_t nK,
                                                      const double* __restrict__ G, int gdim, int nchunk,
                                                      int64_t cbeg, int64_t Ls,
                                                      const void* const* __restrict__ act = nullptr,
                                                      const double* const* __restrict__ X_t = nullptr,
                                                      double* const* __restrict__ E_t = nullptr,
                                                      const double* const* __restrict__ G_t = nullptr) {
    constexpr int NMT = (NK4 + 3) / 4;
    if (act) {
        if (!act[blockIdx.y]) return;
        X = X_t[blockIdx.y];
        E = E_t[blockIdx.y];
        G = G_t[blockIdx.y];
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, cl = lane & 15;
    const int64_t nwork = nK * nchunk;
    const int64_t nblk = (nwork + 7) / 8;
    const int64_t w = xcd_remap(blockIdx.x, nblk) * 8 + wave;
    if (w >= nwork) return;
    const int64_t K = w % nK;
    const int chunk = int(w / nK);
    const uint32_t* ent = clq + K * M;
    uint32_t mb[NK4];
#pragma unroll
    for (int ks = 0; ks < NK4; ++ks) {
        const int m = ks * 4 + g;
        const uint32_t t = ent[m < M ? m : 0];  // unconditional loads (M >= 1): all in flight at once
        mb[ks] = m < M ? t : kInvalid;
    }
    double va[NMT][NK4];
    uint32_t neg[NMT];
#pragma unroll
    for (int mt = 0; mt < NMT; ++mt) {
        const int r = mt * 16 + cl;
        const int pr = int((ent[r < M ? r : 0] >> 21) & 0x3FFu);  // rows r >= M get zero below
#pragma unroll
        for (int ks = 0; ks < NK4; ++ks) {
            const uint32_t e = mb[ks];
            const double t = G[pr * gdim + (e != kInvalid ? int((e >> 21) & 0x3FFu) : 0)];
            const double v = (r < M && e != kInvalid) ? t : 0.0;
            va[mt][ks] = (e >> 31) ? -v : v;  // sign of the gathered row folded into A
        }
        uint32_t nb = 0;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int ro = mt * 16 + 4 * v + g;  // f64 output row
            const uint32_t t = ent[ro < M ? ro : 0];
            nb |= (ro < M ? t >> 31 : 0u) << v;
        }
        neg[mt] = nb;
    }
    const int64_t cend = cbeg + Ls < L ? cbeg + Ls : L;  // even
    const int64_t c0 = cbeg + int64_t(chunk) * CQ3_CW;
    double2 xv[NK4], xn[NK4];
    auto load = [&](int64_t col, double2 (&xx)[NK4]) {
        const int64_t cc = col < cend ? col : cbeg;  // col, cbeg even; past the end: not stored
#pragma unroll
        for (int ks = 0; ks < NK4; ++ks) xx[ks] = *reinterpret_cast<const double2*>(X + gather_row(mb[ks]) * L + cc);
    };
    auto compute = [&](const double2 (&xx)[NK4], int64_t col) {
#pragma unroll
        for (int mt = 0; mt < NMT; ++mt) {
            d4 a0 = d4{0.0, 0.0, 0.0, 0.0}, a1 = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int ks = 0; ks < NK4; ++ks) {
                a0 = __builtin_amdgcn_mfma_f64_16x16x4f64(va[mt][ks], xx[ks].x, a0, 0, 0, 0);
                a1 = __builtin_amdgcn_mfma_f64_16x16x4f64(va[mt][ks], xx[ks].y, a1, 0, 0, 0);
            }
            if (col < cend) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = mt * 16 + 4 * v + g;
                    if (r < M) {
                        const bool ng = (neg[mt] >> v) & 1u;
                        *reinterpret_cast<double2*>(E + ((int64_t(chunk) * nK + K) * M + r) * CQ3_CW + (col - c0)) =
                            ng ? make_double2(-a0[v], -a1[v]) : make_double2(a0[v], a1[v]);
                    }
                }
            }
        }
    };
    // the two 32-column halves of the chunk, the second one's loads in flight during the first
    load(c0 + 2 * cl, xv);
    load(c0 + 32 + 2 * cl, xn);
    compute(xv, c0 + 2 * cl);
    compute(xn, c0 + 32 + 2 * cl);
}

// Pass 2: Y[I][cbeg + c] (= or +=) sum_j E[own[I][j]][c] for c < ncol: every hole string K
// inside I hands I its row of s_m (W_K s X_K)_m. E is either row-major with rows of stride Ls
// (CHUNKED = false) or blocked by 64-column chunk, E[(chunk * nrowE + row) * 64 + c % 64]
// (CHUNKED = true: each pass-1 work item writes one contiguous block). No atomics, and the sum
// runs in a fixed order, so the result does not depend on scheduling.
template <class T, bool ACC, bool CHUNKED>
__global__ __launch_bounds__(256) void k_clique_gather(const T* __restrict__ E, T* __restrict__ Y, int64_t nrow,
                                                       int64_t L, const uint32_t* __restrict__ own, int nown,
                                                       int64_t cbeg, int64_t ncol, int64_t Ls,
                                                       const T* __restrict__ X, const float* __restrict__ gdiag,
                                                       int64_t nrowE,
                                                       const void* const* __restrict__ act = nullptr,
                                                       const T* const* __restrict__ E_t = nullptr,
                                                       T* const* __restrict__ Y_t = nullptr) {
    if (act) {  // batched (FP64 only: no diagonal term)
        if (!act[blockIdx.y]) return;
        E = E_t[blockIdx.y];
        Y = Y_t[blockIdx.y];
    }
    // block: 4 rows x 64 columns
    const int64_t nct = (ncol + 63) / 64;
    const int64_t nwork = ((nrow + 3) / 4) * nct;
    const int64_t w = xcd_remap(blockIdx.x, nwork);
    if (w >= nwork) return;
    const int64_t rb = w / nct, ct = w % nct;
    const int64_t I = rb * 4 + __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    const int64_t cl = ct * 64 + (threadIdx.x & 63);
    if (I >= nrow || cl >= ncol) return;
    const uint32_t* o = own + I * nown;
    // element (row, cl) of E
    const T* Eb = CHUNKED ? E + (cl / 64) * nrowE * 64 + (cl % 64) : E + cl;
    const int64_t rs = CHUNKED ? 64 : Ls;
    T acc = T(0);
    constexpr int GU = 32;  // up to 32 independent loads in flight; the sum runs in j order
    for (int j = 0; j < nown; j += GU) {
        T a[GU];
#pragma unroll
        for (int u = 0; u < GU; ++u) a[u] = j + u < nown ? Eb[int64_t(o[j + u]) * rs] : T(0);
#pragma unroll
        for (int u = 0; u < GU; ++u) acc += a[u];  // + 0 for absent terms is exact
    }
    if (gdiag) acc += T(gdiag[I]) * X[I * L + cbeg + cl];
    T* y = Y + I * L + cbeg + cl;
    if constexpr (ACC)
        *y += acc;
    else
        *y = acc;
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
    double* xd = Xs + KP * AB_TW;                       // [AB_TW] row Ja of C (diagonal term)
    double* srow = xd + AB_TW;                          // [Nb] when use_lds_row
    __shared__ uint32_t ent[AB_MAXE];
    const int64_t Ja = xcd_remap(blockIdx.x, Na);
    if (Ja >= Na) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < ns; i += 256) ent[i] = into_a[Ja * ns + i];
    if (use_lds_row)
        for (int64_t i = tid; i < Nb; i += 256) srow[i] = 0.0;
    __syncthreads();
    const uint32_t occ = str_a[Ja];
    // A operand (V) in registers for this wave's row tiles: V[rs][e] = (p_e q_e | rs)
    double va[AB_MTW][KS];
    double vd[AB_MTW][4];  // diagonal excitations for this lane's output rows
#pragma unroll
    for (int j = 0; j < AB_MTW; ++j) {
        const int mtile = wave + 4 * j;
        const int rs = mtile * 16 + (lane & 15);
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const int e = ks * 4 + (lane >> 4);
            double v = 0.0;
            if (mtile < nmt && rs < np && e < ns) v = vpk[size_t((ent[e] >> 21) & 0x3FFu) * np + rs];
            va[j][ks] = v;
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = mtile * 16 + 4 * v + (lane >> 4);  // f64 output row
            double d = 0.0;
            if (mtile < nmt && r < np) {
                uint32_t o = occ;
                while (o) {
                    int q = __ffs(o) - 1;
                    o &= o - 1u;
                    d += vpk[size_t(q * (q + 1) / 2 + q) * np + r];
                }
            }
            vd[j][v] = d;
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
            if (c0 + c < Nb && e < ns) {
                uint32_t t = ent[e];
                v = C[int64_t(t & kAddrMask) * Nb + c0 + c];
                if (t >> 31) v = -v;
            }
            Xs[e * AB_TW + c] = v;
        }
        for (int c = tid; c < AB_TW; c += 256) xd[c] = (c0 + c < Nb) ? C[Ja * Nb + c0 + c] : 0.0;
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
                const double cd = xd[nt * 16 + (lane & 15)];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const uint32_t t = tgt[j][nt][v];
                    if (t == kInvalid) continue;
                    const double a = acc[nt][v] + vd[j][v] * cd;
                    double y = (t >> 31) ? -a : a;
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
    float* xd = reinterpret_cast<float*>(Xl + AB_TW * KPP);  // [AB_TW] row Ja of C (diagonal term)
    float* srow = xd + AB_TW;
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
    float vd[AB_MTW][4];  // diagonal excitations, sum_{q in Ja} (qq|rs), for this lane's output rows
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
                if (mtile < nmt && rs < np && e < ns) v = vpk[size_t((ent[e] >> 21) & 0x3FFu) * np + rs];
                __bf16 h, l;
                split_bf16(v, h, l);
                vh[j][ks][x] = h;
                vl[j][ks][x] = l;
            }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = mtile * 16 + 4 * (lane >> 4) + v;
            float d = 0.f;
            if (mtile < nmt && r < np) {
                uint32_t o = occ;
                while (o) {
                    int q = __ffs(o) - 1;
                    o &= o - 1u;
                    d += vpk[size_t(q * (q + 1) / 2 + q) * np + r];
                }
            }
            vd[j][v] = d;
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
            if (c0 + c < Nb && e < ns) {
                uint32_t t = ent[e];
                v = C[int64_t(t & kAddrMask) * Nb + c0 + c];
                if (t >> 31) v = -v;
            }
            split_bf16(v, Xh[c * KPP + e], Xl[c * KPP + e]);
        }
        for (int c = tid; c < AB_TW; c += 256) xd[c] = (c0 + c < Nb) ? C[Ja * Nb + c0 + c] : 0.f;
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
                const float cd = xd[nt * 16 + (lane & 15)];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const uint32_t t = tgt[j][nt][v];
                    if (t == kInvalid) continue;
                    const float a = acc[nt][v] + vd[j][v] * cd;
                    float y = (t >> 31) ? -a : a;
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

// ---- alpha-beta, register-resident variant ----------------------------------------------------
// One block per alpha string Ja, 8 waves = WM row-tile groups x (8 / WM) column groups. Each wave
// keeps the A operand V[rs][e] = (p_e q_e | rs) of its row tiles (mtile = wm + WM * j) in
// registers for the whole block and streams its own 32 beta columns at a time: lane (g, c) loads
// columns 2c and 2c + 1 of every gathered row X[e][Ib] = s_e C[Ia_e][Ib] with one 8- (FP32) or
// 16-byte (FP64) load, so each load instruction reads whole 128-byte lines of 4 rows, straight into
// the MFMA B-operand layout of two column tiles (no LDS staging). The product plus the diagonal
// rank-1 term is scattered into the LDS row srow[Jb], Jb = E_rs Ib, from the paired 16-bit target
// table (8 targets per lane and tile = one 16-byte load; a wave's loads are contiguous).
// The LDS row accumulates in FP64: on gfx950 an LDS FP64 atomic add of a full wave costs about a
// seventh of the FP32 one, which is serialized per lane. Requires Nb even and Nb <= 32767.
constexpr uint16_t kInv16 = 0x7FFFu;

template <int KS, int TPW, int WM, int WC>
__global__ __launch_bounds__(64 * WM * WC) void k_ab2_mixed(const float* __restrict__ C, float* __restrict__ S, int64_t Na,
                                                      int64_t Nb, const uint32_t* __restrict__ into_a, int ns,
                                                      const uint32_t* __restrict__ str_a,
                                                      const uint16_t* __restrict__ outT, int npp, int np,
                                                      const float* __restrict__ vpk, int nmt) {
    extern __shared__ __attribute__((aligned(16))) double srow_m[];  // [Nb]
    double* srow = srow_m;
    __shared__ uint32_t ent[KS * 32];
    __shared__ float vdg[12 * 16];
    constexpr int NTH = 64 * WM * WC;
    const int64_t Ja = xcd_remap(blockIdx.x, Na);
    if (Ja >= Na) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave % WM, wc = wave / WM;
    const int g = lane >> 4, cl = lane & 15;
    for (int i = tid; i < KS * 32; i += NTH) ent[i] = i < ns ? into_a[Ja * ns + i] : kInvalid;
    for (int64_t i = tid; i < Nb; i += NTH) srow[i] = 0.0;
    const uint32_t occ = str_a[Ja];
    for (int r = tid; r < nmt * 16; r += NTH) {  // diagonal excitations: sum_{q in Ja} (qq|r)
        float d = 0.f;
        if (r < np)
            for (uint32_t o = occ; o; o &= o - 1u) {
                const int q = __ffs(o) - 1;
                d += vpk[size_t(q * (q + 1) / 2 + q) * np + r];
            }
        vdg[r] = d;
    }
    __syncthreads();
    bf8 ah[TPW][KS], al[TPW][KS];
    float vd[TPW][4];
#pragma unroll
    for (int j = 0; j < TPW; ++j) {
        const int mtile = wm + WM * j;
        const int rs = mtile * 16 + cl;
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
#pragma unroll
            for (int x = 0; x < 8; ++x) {
                const int e = ks * 32 + 8 * g + x;
                const uint32_t en = ent[e];
                const bool use = mtile < nmt && rs < np && e < ns && en != kInvalid;
                const float t = vpk[use ? size_t((en >> 21) & 0x3FFu) * np + rs : 0];  // unconditional load
                const float v = use ? ((en >> 31) ? -t : t) : 0.f;  // sign of the gathered row folded into A
                __bf16 h, l;
                split_bf16(v, h, l);
                ah[j][ks][x] = h;
                al[j][ks][x] = l;
            }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) vd[j][v] = mtile < nmt ? vdg[mtile * 16 + 4 * g + v] : 0.f;
    }
    const float* Crow = C + Ja * Nb;
    const int ntile = npp >> 4;
    for (int64_t c0 = int64_t(wc) * 32; c0 < Nb; c0 += 32 * WC) {
        const int64_t Ib = c0 + 2 * cl;  // and Ib + 1 (Nb even)
        const bool ok = Ib < Nb;
        const int64_t Ibc = ok ? Ib : 0;  // columns past the end: computed, never scattered
        uint4 tg[TPW];
#pragma unroll
        for (int j = 0; j < TPW; ++j) {
            const int mtile = wm + WM * j;  // the table row c0 / 32 exists for all lanes; tiles past nmt
            const int mtc = mtile < nmt ? mtile : 0;  // are skipped below (wave-uniform)
            tg[j] = *reinterpret_cast<const uint4*>(outT + (((c0 >> 5) * ntile + mtc) * 64 + lane) * 8);
        }
        float2 xv[KS][8];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
#pragma unroll
            for (int xx = 0; xx < 8; ++xx)
                xv[ks][xx] = *reinterpret_cast<const float2*>(C + gather_row(ent[ks * 32 + 8 * g + xx]) * Nb + Ibc);
        const float2 cd = *reinterpret_cast<const float2*>(Crow + Ibc);
        bf8 bh0[KS], bl0[KS], bh1[KS], bl1[KS];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
#pragma unroll
            for (int xx = 0; xx < 8; ++xx) {
                __bf16 h, l;
                split_bf16(xv[ks][xx].x, h, l);
                bh0[ks][xx] = h;
                bl0[ks][xx] = l;
                split_bf16(xv[ks][xx].y, h, l);
                bh1[ks][xx] = h;
                bl1[ks][xx] = l;
            }
#pragma unroll
        for (int j = 0; j < TPW; ++j) {
            if (wm + WM * j >= nmt) continue;  // wave-uniform
            f4 a0 = f4{0.f, 0.f, 0.f, 0.f}, a1 = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                a0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[j][ks], bh0[ks], a0, 0, 0, 0);
                a1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[j][ks], bh1[ks], a1, 0, 0, 0);
                a0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[j][ks], bl0[ks], a0, 0, 0, 0);
                a1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[j][ks], bl1[ks], a1, 0, 0, 0);
                a0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[j][ks], bh0[ks], a0, 0, 0, 0);
                a1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[j][ks], bh1[ks], a1, 0, 0, 0);
            }
            const uint32_t tw[4] = {tg[j].x, tg[j].y, tg[j].z, tg[j].w};  // [p][v] pairs of u16
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const uint32_t tt = (tw[p * 2 + v / 2] >> (16 * (v & 1))) & 0xFFFFu;
                    if (!ok || (tt & 0x7FFFu) == kInv16) continue;
                    const float a = (p ? a1[v] : a0[v]) + vd[j][v] * (p ? cd.y : cd.x);
                    atomicAdd(&srow[tt & 0x7FFFu], double((tt >> 15) ? -a : a));
                }
        }
    }
    __syncthreads();
    float* Srow = S + Ja * Nb;
    for (int64_t i = tid; i < Nb; i += NTH) Srow[i] = float(double(Srow[i]) + srow[i]);
}

// Variant with the A operand in LDS (built once per block, read per MFMA; ≈37 KB next to the
// FP64 row) so that the registers it frees hold a second set of gathered rows and targets: the
// loads of the next 32-column tile are in flight while the current one is multiplied and
// scattered (two buffers used alternately). With WM = 1 every wave covers all row tiles of its
// own columns, so each gathered X tile is loaded once per block instead of once per row-tile
// group. The diagonal rank-1 term reads its row factors from LDS (vdg).
template <int KS, int TPW, int WM, int WC>
__global__ __launch_bounds__(64 * WM * WC) void k_ab3_mixed(const float* __restrict__ C, float* __restrict__ S, int64_t Na,
                                                      int64_t Nb, const uint32_t* __restrict__ into_a, int ns,
                                                      const uint32_t* __restrict__ str_a,
                                                      const uint16_t* __restrict__ outT, int npp, int np,
                                                      const float* __restrict__ vpk, int nmt) {
    extern __shared__ __attribute__((aligned(16))) double srow_m3[];  // [Nb] then A operands
    double* srow = srow_m3;
    bf8* Ash = reinterpret_cast<bf8*>(srow + ((Nb + 1) & ~int64_t(1)));  // [nmt][KS][2][64]
    __shared__ uint32_t ent[KS * 32];
    __shared__ __attribute__((aligned(16))) float vdg[12 * 16];
    constexpr int NTH = 64 * WM * WC, NW = WM * WC;
    const int64_t Ja = xcd_remap(blockIdx.x, Na);
    if (Ja >= Na) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave % WM, wc = wave / WM;
    const int g = lane >> 4, cl = lane & 15;
    for (int i = tid; i < KS * 32; i += NTH) ent[i] = i < ns ? into_a[Ja * ns + i] : kInvalid;
    for (int64_t i = tid; i < Nb; i += NTH) srow[i] = 0.0;
    const uint32_t occ = str_a[Ja];
    for (int r = tid; r < nmt * 16; r += NTH) {  // diagonal excitations: sum_{q in Ja} (qq|r)
        float d = 0.f;
        if (r < np)
            for (uint32_t o = occ; o; o &= o - 1u) {
                const int q = __ffs(o) - 1;
                d += vpk[size_t(q * (q + 1) / 2 + q) * np + r];
            }
        vdg[r] = d;
    }
    __syncthreads();
    // A operand slots (wm, j) built by one wave each
    for (int slot = wave; slot < nmt; slot += NW) {
        const int mtile = slot;
        const int rs = mtile * 16 + cl;
        for (int ks = 0; ks < KS; ++ks) {
            bf8 h8, l8;
#pragma unroll
            for (int x = 0; x < 8; ++x) {
                const int e = ks * 32 + 8 * g + x;
                const uint32_t en = ent[e];
                const bool use = rs < np && e < ns && en != kInvalid;
                const float t = vpk[use ? size_t((en >> 21) & 0x3FFu) * np + rs : 0];
                const float v = use ? ((en >> 31) ? -t : t) : 0.f;
                __bf16 h, l;
                split_bf16(v, h, l);
                h8[x] = h;
                l8[x] = l;
            }
            Ash[((slot * KS + ks) * 2 + 0) * 64 + lane] = h8;
            Ash[((slot * KS + ks) * 2 + 1) * 64 + lane] = l8;
        }
    }
    __syncthreads();
    const float* Crow = C + Ja * Nb;
    const int ntile = npp >> 4;
    auto load = [&](int64_t c0, float2 (&x)[KS][8], uint4 (&t)[TPW], float2& d) {
        const int64_t Ib = c0 + 2 * cl;
        const bool ok = Ib < Nb;
        const int64_t Ibc = ok ? Ib : 0;
        const int64_t crow = c0 < Nb ? (c0 >> 5) : 0;  // the prefetch past the last tile reads row 0
#pragma unroll
        for (int j = 0; j < TPW; ++j) {
            const int mtile = wm + WM * j;
            const int mtc = mtile < nmt ? mtile : 0;  // tiles past nmt are skipped (wave-uniform)
            t[j] = *reinterpret_cast<const uint4*>(outT + ((crow * ntile + mtc) * 64 + lane) * 8);
        }
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
#pragma unroll
            for (int xx = 0; xx < 8; ++xx)
                x[ks][xx] = *reinterpret_cast<const float2*>(C + gather_row(ent[ks * 32 + 8 * g + xx]) * Nb + Ibc);
        d = *reinterpret_cast<const float2*>(Crow + Ibc);
    };
    auto compute = [&](const float2 (&x)[KS][8], const uint4 (&t)[TPW], float2 cd, int64_t c0) {
        const bool ok = c0 + 2 * cl < Nb;
        bf8 bh0[KS], bl0[KS], bh1[KS], bl1[KS];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
#pragma unroll
            for (int xx = 0; xx < 8; ++xx) {
                __bf16 h, l;
                split_bf16(x[ks][xx].x, h, l);
                bh0[ks][xx] = h;
                bl0[ks][xx] = l;
                split_bf16(x[ks][xx].y, h, l);
                bh1[ks][xx] = h;
                bl1[ks][xx] = l;
            }
#pragma unroll
        for (int j = 0; j < TPW; ++j) {
            const int slot = wm + WM * j;
            if (slot >= nmt) continue;  // wave-uniform
            f4 a0 = f4{0.f, 0.f, 0.f, 0.f}, a1 = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                const bf8 ah = Ash[((slot * KS + ks) * 2 + 0) * 64 + lane];
                const bf8 al = Ash[((slot * KS + ks) * 2 + 1) * 64 + lane];
                a0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh0[ks], a0, 0, 0, 0);
                a1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh1[ks], a1, 0, 0, 0);
                a0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl0[ks], a0, 0, 0, 0);
                a1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl1[ks], a1, 0, 0, 0);
                a0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh0[ks], a0, 0, 0, 0);
                a1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh1[ks], a1, 0, 0, 0);
            }
            const float4 vq = *reinterpret_cast<const float4*>(vdg + slot * 16 + 4 * g);  // diagonal term rows
            const float vd4[4] = {vq.x, vq.y, vq.z, vq.w};
            const uint32_t tw[4] = {t[j].x, t[j].y, t[j].z, t[j].w};
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const uint32_t tt = (tw[p * 2 + v / 2] >> (16 * (v & 1))) & 0xFFFFu;
                    if (!ok || (tt & 0x7FFFu) == kInv16) continue;
                    const float a = (p ? a1[v] : a0[v]) + vd4[v] * (p ? cd.y : cd.x);
                    atomicAdd(&srow[tt & 0x7FFFu], double((tt >> 15) ? -a : a));
                }
        }
    };
    float2 xa[KS][8], xb[KS][8];
    uint4 ta[TPW], tb[TPW];
    float2 da, db;
    const int64_t step = 32 * WC;
    load(int64_t(wc) * 32, xa, ta, da);
    for (int64_t c0 = int64_t(wc) * 32; c0 < Nb; c0 += 2 * step) {
        load(c0 + step, xb, tb, db);
        compute(xa, ta, da, c0);
        if (c0 + step >= Nb) break;
        load(c0 + 2 * step, xa, ta, da);
        compute(xb, tb, db, c0 + step);
    }
    __syncthreads();
    float* Srow = S + Ja * Nb;
    for (int64_t i = tid; i < Nb; i += NTH) Srow[i] = float(double(Srow[i]) + srow[i]);
}

// FP64: 16x16x4 MFMA, KS k-steps of 4 single excitations. The A operand rows are loaded in the
// permuted order m -> rs = 4 (m % 4) + m / 4 inside each tile so that the accumulator element v of
// lane (g = lane / 16) is row rs = 4 g + v, the layout of the 16-bit target table.
// SPLIT (nmt = 4 (TPW - 1) + 1): the last row tile is shared by the four row-tile waves, each
// taking a quarter of its k-steps and scattering its partial product, so every wave runs
// 2 (TPW - 1) + KS / 4 MFMA chains' worth of k-steps instead of one wave running TPW full tiles.
template <int KS, int TPW, bool SPLIT>
__global__ __launch_bounds__(512, 2) void k_ab2_f64(const double* __restrict__ C, double* __restrict__ S, int64_t Na,
                                                    int64_t Nb, const uint32_t* __restrict__ into_a, int ns,
                                                    const uint32_t* __restrict__ str_a,
                                                    const uint16_t* __restrict__ outT, int npp, int np,
                                                    const double* __restrict__ vpk, int nmt,
                                                    const void* const* __restrict__ act = nullptr,
                                                    const double* const* __restrict__ C_t = nullptr,
                                                    double* const* __restrict__ S_t = nullptr,
                                                    const double* const* __restrict__ V_t = nullptr) {
    extern __shared__ __attribute__((aligned(16))) double srow64[];  // [Nb]
    if (act) {
        if (!act[blockIdx.y]) return;
        C = C_t[blockIdx.y];
        S = S_t[blockIdx.y];
        vpk = V_t[blockIdx.y];
    }
    __shared__ uint32_t ent[KS * 4];
    __shared__ double vdg[12 * 16];
    constexpr int WM = 4, WC = 2, NTH = 64 * WM * WC;
    const int64_t Ja = xcd_remap(blockIdx.x, Na);
    if (Ja >= Na) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave % WM, wc = wave / WM;
    const int g = lane >> 4, cl = lane & 15;
    for (int i = tid; i < KS * 4; i += NTH) ent[i] = i < ns ? into_a[Ja * ns + i] : kInvalid;
    for (int64_t i = tid; i < Nb; i += NTH) srow64[i] = 0.0;
    const uint32_t occ = str_a[Ja];
    for (int r = tid; r < nmt * 16; r += NTH) {  // diagonal excitations: sum_{q in Ja} (qq|r)
        double d = 0.0;
        if (r < np)
            for (uint32_t o = occ; o; o &= o - 1u) {
                const int q = __ffs(o) - 1;
                d += vpk[size_t(q * (q + 1) / 2 + q) * np + r];
            }
        vdg[r] = d;
    }
    __syncthreads();
    constexpr int TF = SPLIT ? TPW - 1 : TPW;  // full row tiles per wave
    constexpr int KQ = KS / 4;                  // k-steps of the shared tile per wave
    static_assert(!SPLIT || KS % 4 == 0, "shared tile needs KS divisible by 4");
    double va[TF][KS];
    double vd[TF][4];
    double vs[SPLIT ? KQ : 1], vds[4];
    const int mts = WM * (TPW - 1);  // the shared tile
    if constexpr (SPLIT) {
        const int rs = mts * 16 + 4 * (cl % 4) + cl / 4;
#pragma unroll
        for (int q = 0; q < KQ; ++q) {
            const int e = (wm * KQ + q) * 4 + g;
            const uint32_t en = ent[e];
            const bool use = mts < nmt && rs < np && e < ns && en != kInvalid;
            const double t = vpk[use ? size_t((en >> 21) & 0x3FFu) * np + rs : 0];
            vs[q] = use ? ((en >> 31) ? -t : t) : 0.0;
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) vds[v] = wm == 0 && mts < nmt ? vdg[mts * 16 + 4 * g + v] : 0.0;  // once
    }
#pragma unroll
    for (int j = 0; j < TF; ++j) {
        const int mtile = wm + WM * j;
        const int rs = mtile * 16 + 4 * (cl % 4) + cl / 4;  // permuted A row
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const int e = ks * 4 + g;
            const uint32_t en = ent[e];
            const bool use = mtile < nmt && rs < np && e < ns && en != kInvalid;
            const double t = vpk[use ? size_t((en >> 21) & 0x3FFu) * np + rs : 0];  // unconditional load
            va[j][ks] = use ? ((en >> 31) ? -t : t) : 0.0;  // sign of the gathered row folded into A
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) vd[j][v] = mtile < nmt ? vdg[mtile * 16 + 4 * g + v] : 0.0;
    }
    const double* Crow = C + Ja * Nb;
    const int ntile = npp >> 4;
    // the gathered rows, targets and C[Ja] columns of one 32-column tile; the next tile's are
    // loaded after the current tile's products, so they are in flight during its scatter
    auto load = [&](int64_t c0, double2 (&xv)[KS], uint4 (&tg)[TPW], double2& cd) {
        const int64_t Ib = c0 + 2 * cl;
        const int64_t Ibc = Ib < Nb ? Ib : 0;  // columns past the end: computed, never scattered
#pragma unroll
        for (int j = 0; j < TPW; ++j) {
            const int mtile = SPLIT && j == TPW - 1 ? mts : wm + WM * j;  // the table row c0 / 32 exists
            const int mtc = mtile < nmt ? mtile : 0;  // for all lanes; tiles past nmt are skipped below
            tg[j] = *reinterpret_cast<const uint4*>(outT + (((c0 >> 5) * ntile + mtc) * 64 + lane) * 8);
        }
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
            xv[ks] = *reinterpret_cast<const double2*>(C + gather_row(ent[ks * 4 + g]) * Nb + Ibc);
        cd = *reinterpret_cast<const double2*>(Crow + Ibc);
    };
    double2 xv[KS];
    uint4 tg[TPW];
    double2 cd;
    int64_t c0 = int64_t(wc) * 32;
    if (c0 < Nb) load(c0, xv, tg, cd);
    for (; c0 < Nb; c0 += 32 * WC) {
        const bool ok = c0 + 2 * cl < Nb;
        d4 acc0[TPW], acc1[TPW];
#pragma unroll
        for (int j = 0; j < TF; ++j) {
            acc0[j] = d4{0.0, 0.0, 0.0, 0.0};
            acc1[j] = d4{0.0, 0.0, 0.0, 0.0};
            if (wm + WM * j >= nmt) continue;  // wave-uniform
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                acc0[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(va[j][ks], xv[ks].x, acc0[j], 0, 0, 0);
                acc1[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(va[j][ks], xv[ks].y, acc1[j], 0, 0, 0);
            }
        }
        if constexpr (SPLIT) {
            acc0[TPW - 1] = d4{0.0, 0.0, 0.0, 0.0};
            acc1[TPW - 1] = d4{0.0, 0.0, 0.0, 0.0};
            if (mts < nmt) {  // this wave's quarter of the shared tile's k-steps (compile-time indices)
#pragma unroll
                for (int w = 0; w < WM; ++w)
                    if (wm == w) {
#pragma unroll
                        for (int q = 0; q < KQ; ++q) {
                            acc0[TPW - 1] =
                                __builtin_amdgcn_mfma_f64_16x16x4f64(vs[q], xv[w * KQ + q].x, acc0[TPW - 1], 0, 0, 0);
                            acc1[TPW - 1] =
                                __builtin_amdgcn_mfma_f64_16x16x4f64(vs[q], xv[w * KQ + q].y, acc1[TPW - 1], 0, 0, 0);
                        }
                    }
            }
        }
        uint4 tgc[TPW];
#pragma unroll
        for (int j = 0; j < TPW; ++j) tgc[j] = tg[j];
        const double2 cdc = cd;
        if (c0 + 32 * WC < Nb) load(c0 + 32 * WC, xv, tg, cd);  // wave-uniform
        auto scatter = [&](const uint4& t4, const d4& a0, const d4& a1, const double (&vdj)[4]) {
            const uint32_t tw[4] = {t4.x, t4.y, t4.z, t4.w};
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const uint32_t tt = (tw[p * 2 + v / 2] >> (16 * (v & 1))) & 0xFFFFu;
                    if (!ok || (tt & 0x7FFFu) == kInv16) continue;
                    const double a = (p ? a1[v] : a0[v]) + vdj[v] * (p ? cdc.y : cdc.x);
                    atomicAdd(&srow64[tt & 0x7FFFu], (tt >> 15) ? -a : a);
                }
        };
#pragma unroll
        for (int j = 0; j < TF; ++j) {
            if (wm + WM * j >= nmt) continue;  // wave-uniform
            scatter(tgc[j], acc0[j], acc1[j], vd[j]);
        }
        if constexpr (SPLIT)
            if (mts < nmt) scatter(tgc[TPW - 1], acc0[TPW - 1], acc1[TPW - 1], vds);
    }
    __syncthreads();
    double* Srow = S + Ja * Nb;
    for (int64_t i = tid; i < Nb; i += NTH) Srow[i] += srow64[i];
}

__global__ void k_clique_diag(const uint32_t* __restrict__ str, int64_t nstr, int n, int d,
                              const double* __restrict__ g, int gdim, float* __restrict__ out) {
    const int64_t I = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (I >= nstr) return;
    const uint32_t s = str[I];
    double acc = 0.0;
    for (int a = 0; a < n; ++a) {
        if (!((s >> a) & 1u)) continue;
        if (d == 1) {
            acc += g[size_t(a) * gdim + a];
            continue;
        }
        for (int b = 0; b < a; ++b) {
            if (!((s >> b) & 1u)) continue;
            const int P = a * (a - 1) / 2 + b;
            acc += g[size_t(P) * gdim + P];
        }
    }
    out[I] = float(acc);
}

void clique_diagonal_device(const SpinTables& sp, const double* d_g, int gdim, float* d_out, hipStream_t st) {
    k_clique_diag<<<nblk(sp.nstr, 128), 128, 0, st>>>(sp.str.p, sp.nstr, sp.n, sp.d, d_g, gdim, d_out);
    HIPCHK(hipGetLastError());
}

// ---- launchers -----------------------------------------------------------------------------
struct SigmaWork {
    DBuf<double> t64a, t64b;
    DBuf<float> t32a, t32b;
    DBuf<unsigned char> clique;  // two-pass same-spin scratch
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

// Bytes of scratch for the two-pass same-spin product (E holds nK*M rows of a column slab). The
// slab is sized so that E stays in the 256 MB memory-side cache between pass 1 and pass 2: 192 MB
// measured fastest at CAS(16,16) in both precisions (128-column slabs in FP32, 64 in FP64).
static size_t clique_scratch_bytes() {
    static size_t v = [] {
        size_t mb = 192;
        if (const char* e = std::getenv("FCI_CLIQUE_SCRATCH_MB")) {
            long x = std::atol(e);
            if (x >= 16) mb = size_t(x);
        }
        return mb << 20;
    }();
    return v;
}

// Same-spin product Y (+)= sum_K P_K^T S_K W_K S_K P_K X over the pair cliques of one spin, X and
// Y with L columns. Pass 1 (k_clique2_*<STORE>) writes E_K = S_K W_K S_K X_K for a column chunk,
// pass 2 (k_clique_gather) sums, for every string, the C(k,d) rows of E that address it.
// True when launch_clique runs the two-pass path (and so needs the intermediate E).
template <class T> static bool clique_two_pass(const SpinTables& sp, const DevInts& ints, int spin) {
    if (sp.d == 0 || sp.nK == 0) return false;
    const int gd = ints.gdim[spin];
    const int Mp = (sp.M + 15) & ~15, Mk = (sp.M + 31) & ~31;
    const size_t sh = sizeof(T) == 8 ? size_t(8) * Mp * 16 * 8 + size_t(8) * Mp * 4
                                     : ((size_t(gd) * gd + 3) & ~size_t(3)) * 4 +
                                           size_t(8) * 2 * 32 * (Mk + 8) * 2 + size_t(8) * Mk * 4;
    return !(sh > 150 * 1024 || sp.M > 1023);
}
// Column width of one pass-1 work item (the slab width is a multiple of it).
template <class T> static int clique_tile_width(const SpinTables& sp) {
    return sp.M <= 64 ? CQ3_CW : (sizeof(T) == 8 ? 16 : 32);
}

// ext != nullptr: E comes from the caller's ext_bytes instead of SigmaWork (never allocated).
template <class T>
static void launch_clique(const T* X, T* Y, int64_t L, const SpinTables& sp, const DevInts& ints, int spin,
                          bool accumulate, SigmaWork& w, hipStream_t st, unsigned char* ext = nullptr,
                          size_t ext_bytes = 0) {
    if (sp.d == 0 || sp.nK == 0) {
        if (!accumulate) k_fill<T><<<nblk(sp.nstr * L, 256), 256, 0, st>>>(Y, sp.nstr * L, T(0));
        return;
    }
    const int gd = ints.gdim[spin];
    const int Mp = (sp.M + 15) & ~15, Mk = (sp.M + 31) & ~31;
    constexpr int TW64 = 16, TW32 = 32, WV64 = 8, WV32 = 8;
    const size_t sh = sizeof(T) == 8 ? size_t(WV64) * Mp * TW64 * 8 + size_t(WV64) * Mp * 4
                                     : ((size_t(gd) * gd + 3) & ~size_t(3)) * 4 +
                                           size_t(WV32) * 2 * TW32 * (Mk + 8) * 2 + size_t(WV32) * Mk * 4;
    if (!clique_two_pass<T>(sp, ints, spin)) {
        // large cliques: one-pass kernel with global atomics
        if (sp.M > CQ_MAXM) throw Error(-1, "clique size exceeds kernel limit");
        if (!accumulate) k_fill<T><<<nblk(sp.nstr * L, 256), 256, 0, st>>>(Y, sp.nstr * L, T(0));
        int nct = int((L + CQ_TW - 1) / CQ_TW), nmt = (sp.M + CQ_MT - 1) / CQ_MT;
        int64_t nwork = int64_t(nct) * sp.nK * nmt;
        if constexpr (sizeof(T) == 8)
            k_clique_f64<<<xcd_grid(nwork), 256, 0, st>>>(X, Y, L, sp.clq.p, sp.M, sp.nK, ints.g2[spin].p,
                                                          ints.gdim[spin], nct, nmt);
        else
            k_clique_mixed<<<xcd_grid(nwork), 256, 0, st>>>(X, Y, L, sp.clq.p, sp.M, sp.nK, ints.g2f[spin].p,
                                                            ints.gdim[spin], nct, nmt);
        HIPCHK(hipGetLastError());
        return;
    }
    const bool reg = sp.M <= 64;  // register-resident pass 1
    const int tw = clique_tile_width<T>(sp);
    const int wv = sizeof(T) == 8 ? WV64 : WV32;
    const int64_t rowsE = sp.nK * sp.M;
    const int64_t Lr = (L + tw - 1) / tw * tw;
    const size_t cap = ext ? std::min(ext_bytes, clique_scratch_bytes()) : clique_scratch_bytes();
    int64_t Ls = int64_t(cap / (size_t(rowsE) * sizeof(T))) / tw * tw;
    Ls = std::max<int64_t>(tw, std::min<int64_t>(Ls, Lr));
    const size_t need = size_t(rowsE) * size_t(Ls) * sizeof(T);
    T* E;
    if (ext) {
        if (need > ext_bytes) throw Error(-1, "sigma: external scratch too small for the same-spin intermediate");
        E = reinterpret_cast<T*>(ext);
    } else {
        DBuf<unsigned char>& eb = w.clique;
        if (eb.n < need) {
            HIPCHK(hipStreamSynchronize(st));
            eb.alloc(need);
        }
        E = reinterpret_cast<T*>(eb.p);
    }
    const int nct = int(Ls / tw);
    int kch = 64;
    while (kch > wv && int64_t(nct) * ((sp.nK + kch - 1) / kch) < 2048) kch /= 2;
    const int nkc = int((sp.nK + kch - 1) / kch);
    const int64_t nwork = int64_t(nct) * nkc;
    if constexpr (sizeof(T) == 8)
        HIPCHK(hipFuncSetAttribute((const void*)k_clique2_f64<TW64, WV64, true>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, int(sh)));
    else
        HIPCHK(hipFuncSetAttribute((const void*)k_clique2_mixed<TW32, WV32, true>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, int(sh)));
    for (int64_t cbeg = 0; cbeg < L; cbeg += Ls) {
        const int64_t ncol = std::min<int64_t>(Ls, L - cbeg);
        if (reg) {
            const int nchunk = int((ncol + CQ3_CW - 1) / CQ3_CW);
            // g' stays in global memory (L1/L2): staging it in LDS caps the resident blocks per CU
            // and measured slower at CAS(16,16) in both precisions
            const size_t gsh = size_t(gd) * gd * sizeof(T);
            const bool gl = false;
            static const bool p64 = [] {
                const char* e = std::getenv("FCI_CQ_P64");
                return !e || std::atoi(e) != 0;
            }();
            if constexpr (sizeof(T) == 8) {
                const int64_t nblk3 = (sp.nK * nchunk + 7) / 8;
                const int nk4 = (sp.M + 3) / 4;
                if (p64 && L % 2 == 0 && nk4 <= 16) {  // paired columns
                    switch (nk4) {
#define C3P(K)                                                                                            \
    case K:                                                                                               \
        k_clique3p_f64<K><<<xcd_grid(nblk3), 512, 0, st>>>(X, E, L, sp.clq.p, sp.M, sp.nK, ints.g2[spin].p, \
                                                           gd, nchunk, cbeg, Ls);                         \
        break;
                        C3P(1) C3P(2) C3P(3) C3P(4) C3P(5) C3P(6) C3P(7) C3P(8) C3P(9) C3P(10) C3P(11) C3P(12) C3P(13)
                        C3P(14) C3P(15) C3P(16)
#undef C3P
                        default: throw Error(-1, "clique kernel: no instance for this size");
                    }
                } else
                switch (nk4) {
#define C3(K)                                                                                                   \
    case K:                                                                                                     \
        if (gl) {                                                                                               \
            HIPCHK(hipFuncSetAttribute((const void*)k_clique3_f64<K, true>,                                    \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(gsh)));                  \
            k_clique3_f64<K, true><<<xcd_grid(nblk3), 512, gsh, st>>>(X, E, L, sp.clq.p, sp.M, sp.nK,           \
                                                                     ints.g2[spin].p, gd, nchunk, cbeg, Ls);   \
        } else {                                                                                                \
            k_clique3_f64<K, false><<<xcd_grid(nblk3), 512, 0, st>>>(X, E, L, sp.clq.p, sp.M, sp.nK,            \
                                                                    ints.g2[spin].p, gd, nchunk, cbeg, Ls);    \
        }                                                                                                       \
        break;
                    C3(1) C3(2) C3(3) C3(4) C3(5) C3(6) C3(7) C3(8) C3(9) C3(10) C3(11) C3(12) C3(13) C3(14) C3(15) C3(16)
#undef C3
                    default: throw Error(-1, "clique kernel: no instance for this size");
                }
            } else if (L % 2 == 0) {
                static const int cpw = [] {
                    const char* e = std::getenv("FCI_CQ_CPW");
                    return e ? std::max(1, std::atoi(e)) : 4;
                }();
                const int64_t nblk3p = (sp.nK * ((nchunk + cpw - 1) / cpw) + 3) / 4;
                const int nmt = (sp.M + 15) / 16;
                switch (nmt) {
#define C3(K)                                                                                                 \
    case K:                                                                                                   \
        if (sp.M == 28 && K == 2)                                                                             \
            k_clique3p_mixed<K, 28><<<xcd_grid(nblk3p), 256, 0, st>>>(X, E, L, sp.clq.p, sp.M, sp.nK,         \
                                                                      ints.g2f_off[spin].p, gd, nchunk, cbeg, Ls, cpw); \
        else if (sp.M == 36 && K == 3)                                                                        \
            k_clique3p_mixed<K, 36><<<xcd_grid(nblk3p), 256, 0, st>>>(X, E, L, sp.clq.p, sp.M, sp.nK,         \
                                                                      ints.g2f_off[spin].p, gd, nchunk, cbeg, Ls, cpw); \
        else if (sp.M == 45 && K == 3)                                                                        \
            k_clique3p_mixed<K, 45><<<xcd_grid(nblk3p), 256, 0, st>>>(X, E, L, sp.clq.p, sp.M, sp.nK,         \
                                                                      ints.g2f_off[spin].p, gd, nchunk, cbeg, Ls, cpw); \
        else if (sp.M == 55 && K == 4)                                                                        \
            k_clique3p_mixed<K, 55><<<xcd_grid(nblk3p), 256, 0, st>>>(X, E, L, sp.clq.p, sp.M, sp.nK,         \
                                                                      ints.g2f_off[spin].p, gd, nchunk, cbeg, Ls, cpw); \
        else                                                                                                  \
            k_clique3p_mixed<K, 0><<<xcd_grid(nblk3p), 256, 0, st>>>(X, E, L, sp.clq.p, sp.M, sp.nK,          \
                                                                     ints.g2f_off[spin].p, gd, nchunk, cbeg, Ls, cpw); \
        break;
                    C3(1) C3(2) C3(3) C3(4)
#undef C3
                    default: throw Error(-1, "clique kernel: no instance for this size");
                }
            } else {
                const int64_t nblk3 = (sp.nK * nchunk + 3) / 4;
                const int nmt = (sp.M + 15) / 16;
                switch (nmt) {
#define C3(K)                                                                                                   \
    case K:                                                                                                     \
        if (gl) {                                                                                               \
            HIPCHK(hipFuncSetAttribute((const void*)k_clique3_mixed<K, true>,                                  \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(gsh)));                  \
            k_clique3_mixed<K, true><<<xcd_grid(nblk3), 256, gsh, st>>>(X, E, L, sp.clq.p, sp.M, sp.nK,         \
                                                                       ints.g2f_off[spin].p, gd, nchunk, cbeg, Ls); \
        } else {                                                                                                \
            k_clique3_mixed<K, false><<<xcd_grid(nblk3), 256, 0, st>>>(X, E, L, sp.clq.p, sp.M, sp.nK,          \
                                                                      ints.g2f_off[spin].p, gd, nchunk, cbeg, Ls);  \
        }                                                                                                       \
        break;
                    C3(1) C3(2) C3(3) C3(4)
#undef C3
                    default: throw Error(-1, "clique kernel: no instance for this size");
                }
            }
        } else if constexpr (sizeof(T) == 8)
            k_clique2_f64<TW64, WV64, true><<<xcd_grid(nwork), 64 * WV64, sh, st>>>(
                X, E, L, sp.clq.p, sp.M, sp.nK, ints.g2[spin].p, gd, nct, nkc, kch, cbeg, Ls);
        else
            k_clique2_mixed<TW32, WV32, true><<<xcd_grid(nwork), 64 * WV32, sh, st>>>(
                X, E, L, sp.clq.p, sp.M, sp.nK, ints.g2f_off[spin].p, gd, nct, nkc, kch, cbeg, Ls);
        HIPCHK(hipGetLastError());
        const int64_t gw = ((sp.nstr + 3) / 4) * ((ncol + 63) / 64);
        // mixed: the operator diagonal was left out of pass 1 and is applied here in FP32
        const float* gdiag = sizeof(T) == 4 ? ints.gdiag[spin].p : nullptr;
        const int64_t nrowE = sp.nK * sp.M;
        if (accumulate && reg)
            k_clique_gather<T, true, true><<<xcd_grid(gw), 256, 0, st>>>(E, Y, sp.nstr, L, sp.own.p, sp.nown,
                                                                         cbeg, ncol, Ls, X, gdiag, nrowE);
        else if (reg)
            k_clique_gather<T, false, true><<<xcd_grid(gw), 256, 0, st>>>(E, Y, sp.nstr, L, sp.own.p, sp.nown,
                                                                          cbeg, ncol, Ls, X, gdiag, nrowE);
        else if (accumulate)
            k_clique_gather<T, true, false><<<xcd_grid(gw), 256, 0, st>>>(E, Y, sp.nstr, L, sp.own.p, sp.nown,
                                                                          cbeg, ncol, Ls, X, gdiag, nrowE);
        else
            k_clique_gather<T, false, false><<<xcd_grid(gw), 256, 0, st>>>(E, Y, sp.nstr, L, sp.own.p, sp.nown,
                                                                           cbeg, ncol, Ls, X, gdiag, nrowE);
        HIPCHK(hipGetLastError());
    }
}

template <int KS>
static void ab_f64_launch(const double* C, double* S, const SpinTables& a, const SpinTables& b,
                          const DevInts& ints, hipStream_t st, bool half) {
    int nmt = (b.np + 15) / 16;
    size_t xs = size_t(KS * 4) * AB_TW * sizeof(double) + AB_TW * sizeof(double);
    size_t row = size_t(b.nstr) * sizeof(double);
    int use_row = (xs + row) <= 150 * 1024;
    size_t sh = xs + (use_row ? row : 0);
    auto kern = k_ab_f64<KS>;
    HIPCHK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, int(sh)));
    kern<<<xcd_grid(a.nstr), 256, sh, st>>>(C, S, a.nstr, b.nstr, a.into.p, a.ns, a.str.p, b.out.p, b.np,
                                            (half ? ints.vpk_h.p : ints.vpk.p), nmt, use_row);
    HIPCHK(hipGetLastError());
}
template <int KS>
static void ab_mixed_launch(const float* C, float* S, const SpinTables& a, const SpinTables& b,
                            const DevInts& ints, hipStream_t st, bool half) {
    int nmt = (b.np + 15) / 16;
    size_t xs = size_t(AB_TW) * (KS * 32 + 8) * 2 * sizeof(__bf16) + AB_TW * sizeof(float);
    size_t row = size_t(b.nstr) * sizeof(float);
    int use_row = (xs + row) <= 150 * 1024;
    size_t sh = xs + (use_row ? row : 0);
    auto kern = k_ab_mixed<KS>;
    HIPCHK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, int(sh)));
    kern<<<xcd_grid(a.nstr), 256, sh, st>>>(C, S, a.nstr, b.nstr, a.into.p, a.ns, a.str.p, b.out.p, b.np,
                                            (half ? ints.vpk32_h.p : ints.vpk32.p), nmt, use_row);
    HIPCHK(hipGetLastError());
}

template <int KS, int TPW, bool SPLIT>
static void ab2_f64_launch(const double* C, double* S, const SpinTables& a, const SpinTables& b,
                           const DevInts& ints, hipStream_t st, bool half) {
    const size_t sh = size_t(b.nstr) * sizeof(double);
    auto kern = k_ab2_f64<KS, TPW, SPLIT>;
    HIPCHK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, int(sh)));
    kern<<<xcd_grid(a.nstr), 512, sh, st>>>(C, S, a.nstr, b.nstr, a.into.p, a.ns, a.str.p, b.outT.p, b.npp, b.np,
                                            (half ? ints.vpk_h.p : ints.vpk.p), (b.np + 15) / 16, nullptr, nullptr, nullptr, nullptr);
    HIPCHK(hipGetLastError());
}
template <int KS, int TPW, int WM, int WC>
static void ab3_mixed_launch(const float* C, float* S, const SpinTables& a, const SpinTables& b,
                             const DevInts& ints, hipStream_t st, bool half) {
    const int nmt = (b.np + 15) / 16;
    const size_t sh3 = size_t((b.nstr + 1) & ~int64_t(1)) * sizeof(double) + size_t(nmt) * KS * 2 * 64 * 16;
    if (nmt > WM * TPW || sh3 + KS * 32 * 4 + 12 * 16 * 4 > 160 * 1024)
        throw Error(-1, "alpha-beta kernel: LDS A-operand layout does not fit this shape");
    auto kern3 = k_ab3_mixed<KS, TPW, WM, WC>;
    HIPCHK(hipFuncSetAttribute((const void*)kern3, hipFuncAttributeMaxDynamicSharedMemorySize, int(sh3)));
    kern3<<<xcd_grid(a.nstr), 64 * WM * WC, sh3, st>>>(C, S, a.nstr, b.nstr, a.into.p, a.ns, a.str.p, b.outT.p,
                                                       b.npp, b.np, (half ? ints.vpk32_h.p : ints.vpk32.p), nmt);
    HIPCHK(hipGetLastError());
}

template <int KS, int TPW, int WM, int WC>
static void ab2_mixed_launch(const float* C, float* S, const SpinTables& a, const SpinTables& b,
                             const DevInts& ints, hipStream_t st, bool half) {
    const size_t sh = size_t(b.nstr) * sizeof(double);
    auto kern = k_ab2_mixed<KS, TPW, WM, WC>;
    HIPCHK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, int(sh)));
    kern<<<xcd_grid(a.nstr), 64 * WM * WC, sh, st>>>(C, S, a.nstr, b.nstr, a.into.p, a.ns, a.str.p, b.outT.p, b.npp, b.np,
                                            (half ? ints.vpk32_h.p : ints.vpk32.p), (b.np + 15) / 16);
    HIPCHK(hipGetLastError());
}

// register-resident kernels: 16-bit targets, LDS row, E <= 96, <= 12 row tiles
static bool ab2_ok(const SpinTables& a, const SpinTables& b, size_t w) {
    return b.outT.p != nullptr && b.nstr % 2 == 0 && a.ns <= 96 && (b.np + 15) / 16 <= 12 &&
           size_t(b.nstr) * w <= 150 * 1024;
}

static void launch_ab(const double* C, double* S, const SpinTables& a, const SpinTables& b,
                      const DevInts& ints, hipStream_t st, bool half) {
    if (ab2_ok(a, b, sizeof(double))) {
        const int ks = std::max(1, (a.ns + 15) / 16) * 4;  // k-steps of 4, in buckets of 16 excitations
        const int tpw = ((b.np + 15) / 16 + 3) / 4;
        // nmt = 4 (tpw - 1) + 1: the last row tile shared by the four row-tile waves
        const bool split = tpw > 1 && (b.np + 15) / 16 == 4 * (tpw - 1) + 1;
#define F64CASE(K, T)                                                                      \
    if (ks == K && tpw == T)                                                                \
        return split ? ab2_f64_launch<K, T, (T > 1)>(C, S, a, b, ints, st, half)                  \
                     : ab2_f64_launch<K, T, false>(C, S, a, b, ints, st, half);
#define F64ROW(K) F64CASE(K, 1) F64CASE(K, 2) F64CASE(K, 3)
        F64ROW(4) F64ROW(8) F64ROW(12) F64ROW(16) F64ROW(20) F64ROW(24)
#undef F64ROW
#undef F64CASE
        throw Error(-1, "alpha-beta kernel: no instance for this shape");
    }
    int ks = std::max(1, (a.ns + 3) / 4);
    if ((b.np + 15) / 16 > 4 * AB_MTW) throw Error(-1, "too many orbitals for the alpha-beta kernel");
    switch (ks) {
#define CASE(K) \
    case K: ab_f64_launch<K>(C, S, a, b, ints, st, half); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10)
        CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16) CASE(17) CASE(18) CASE(19) CASE(20)
        CASE(21) CASE(22) CASE(23) CASE(24) CASE(25) CASE(26) CASE(27) CASE(28) CASE(29) CASE(30)
        CASE(31) CASE(32) CASE(33) CASE(34) CASE(35) CASE(36) CASE(37) CASE(38) CASE(39) CASE(40)
#undef CASE
        default: throw Error(-1, "alpha-beta kernel: too many single excitations");
    }
}
static void launch_ab(const float* C, float* S, const SpinTables& a, const SpinTables& b,
                      const DevInts& ints, hipStream_t st, bool half) {
    if (ab2_ok(a, b, sizeof(double))) {
        const int ks = std::max(1, (a.ns + 31) / 32);
        const int nmt = (b.np + 15) / 16;
        // Default for 3..10 row tiles: 8 waves each on its own 32 columns and all row tiles, the A
        // operand in LDS, the next tile's loads in flight (k_ab3_mixed<K, T, 1, 8>: 16.5 ms at
        // CAS(16,16)); otherwise 8 waves as 2 row-tile groups x 4 column groups with A in
        // registers (k_ab2_mixed<K, T, 2, 4>: 19.7 ms there, each X tile loaded by two waves).
        // FCI_AB_LAYOUT=8 forces the latter, 12 and 3 select 4 x 3 and 3 x 4 groups.
        static const int layout = [] {
            const char* e = std::getenv("FCI_AB_LAYOUT");
            return e ? std::atoi(e) : 1;
        }();
        if (layout == 3 && (nmt + 2) / 3 * ks <= 6) {
            const int tpw = (nmt + 2) / 3;
#define MXCASE(K, T) \
    if (ks == K && tpw == T) return ab2_mixed_launch<K, T, 3, 4>(C, S, a, b, ints, st, half);
            MXCASE(1, 1) MXCASE(1, 2) MXCASE(1, 3) MXCASE(1, 4) MXCASE(2, 1) MXCASE(2, 2) MXCASE(2, 3)
            MXCASE(3, 1) MXCASE(3, 2)
#undef MXCASE
        }
        if (layout == 12 && (nmt + 3) / 4 * ks <= 6) {
            const int tpw = (nmt + 3) / 4;
#define MXCASE(K, T) \
    if (ks == K && tpw == T) return ab2_mixed_launch<K, T, 4, 3>(C, S, a, b, ints, st, half);
            MXCASE(1, 1) MXCASE(1, 2) MXCASE(1, 3) MXCASE(2, 1) MXCASE(2, 2) MXCASE(2, 3)
            MXCASE(3, 1) MXCASE(3, 2)
#undef MXCASE
        }
        const size_t sh3 = size_t((b.nstr + 1) & ~int64_t(1)) * 8 + size_t(nmt) * ks * 2 * 64 * 16 + 1024 + 768;
        if (layout == 1 && sh3 <= 160 * 1024) {
#define MXCASE(K, T) \
    if (ks == K && nmt == T) return ab3_mixed_launch<K, T, 1, 8>(C, S, a, b, ints, st, half);
            MXCASE(1, 3) MXCASE(1, 4) MXCASE(1, 5) MXCASE(1, 6) MXCASE(1, 7) MXCASE(1, 8) MXCASE(1, 9) MXCASE(1, 10)
            MXCASE(2, 3) MXCASE(2, 4) MXCASE(2, 5) MXCASE(2, 6) MXCASE(2, 7) MXCASE(2, 8) MXCASE(2, 9) MXCASE(2, 10)
#undef MXCASE
        }
        const int tpw = (nmt + 1) / 2;
#define MXCASE(K, T) \
    if (ks == K && tpw == T) return ab2_mixed_launch<K, T, 2, 4>(C, S, a, b, ints, st, half);
#define MXROW(K) MXCASE(K, 1) MXCASE(K, 2) MXCASE(K, 3) MXCASE(K, 4) MXCASE(K, 5) MXCASE(K, 6)
        MXROW(1) MXROW(2) MXROW(3)
#undef MXROW
#undef MXCASE
        throw Error(-1, "alpha-beta kernel: no instance for this shape");
    }
    int ks = std::max(1, (a.ns + 31) / 32);
    if ((b.np + 15) / 16 > 4 * AB_MTW) throw Error(-1, "too many orbitals for the alpha-beta kernel");
    switch (ks) {
        case 1: ab_mixed_launch<1>(C, S, a, b, ints, st, half); break;
        case 2: ab_mixed_launch<2>(C, S, a, b, ints, st, half); break;
        case 3: ab_mixed_launch<3>(C, S, a, b, ints, st, half); break;
        case 4: ab_mixed_launch<4>(C, S, a, b, ints, st, half); break;
        case 5: ab_mixed_launch<5>(C, S, a, b, ints, st, half); break;
        default: throw Error(-1, "alpha-beta kernel: too many single excitations");
    }
}

// ---- batched FP64 sigma (many problems of one size, one launch per term) ------------------------
// Supported: both spins with >= 2 electrons, same-spin cliques of <= 64 members with <= 16 k-steps,
// even string counts (paired-column kernels) and the register-resident alpha-beta kernel.
bool batch_sigma_supported(const SpinTables& a, const SpinTables& b, std::string* why) {
    auto no = [&](const char* m) {
        if (why) *why = m;
        return false;
    };
    if (a.d != 2 || b.d != 2) return no("batched sigma needs at least 2 electrons of each spin");
    if (a.M > 64 || b.M > 64 || (a.M + 3) / 4 > 16 || (b.M + 3) / 4 > 16)
        return no("batched sigma: same-spin cliques too large");
    if (a.nstr % 2 || b.nstr % 2) return no("batched sigma needs even string counts");
    if (!ab2_ok(a, b, sizeof(double))) return no("batched sigma: alpha-beta shape not supported");
    const int ks = std::max(1, (a.ns + 15) / 16) * 4;
    const int tpw = ((b.np + 15) / 16 + 3) / 4;
    if (ks > 24 || tpw > 3) return no("batched sigma: alpha-beta shape not supported");
    return true;
}

static int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

size_t batch_sigma_scratch_elems(const SpinTables& a, const SpinTables& b, size_t* off_yt, size_t* off_e) {
    const size_t nd = size_t(a.nstr) * size_t(b.nstr);
    const size_t v = (nd + 31) & ~size_t(31);  // 256-byte aligned in FP64
    const size_t ea = size_t(a.nK) * a.M * size_t(round_up(b.nstr, CQ3_CW));
    const size_t eb = size_t(b.nK) * b.M * size_t(round_up(a.nstr, CQ3_CW));
    if (off_yt) *off_yt = v;
    if (off_e) *off_e = 2 * v;
    return 2 * v + std::max(ea, eb);
}

static void batch_clique_f64(const SpinTables& sp, int gd, int64_t L, const void* const* act,
                             const double* const* X_t, double* const* E_t, const double* const* G_t,
                             double* const* Y_t, bool acc, int nitem, hipStream_t st) {
    const int64_t Ls = round_up(L, CQ3_CW);
    const int nchunk = int(Ls / CQ3_CW);
    const int64_t nblk3 = (sp.nK * nchunk + 7) / 8;
    const dim3 g1(xcd_grid(nblk3), unsigned(nitem));
    switch ((sp.M + 3) / 4) {
#define C3P(K)                                                                                              \
    case K:                                                                                                 \
        k_clique3p_f64<K><<<g1, 512, 0, st>>>(nullptr, nullptr, L, sp.clq.p, sp.M, sp.nK, nullptr, gd, nchunk, \
                                              0, Ls, act, X_t, E_t, G_t);                                   \
        break;
        C3P(1) C3P(2) C3P(3) C3P(4) C3P(5) C3P(6) C3P(7) C3P(8) C3P(9) C3P(10) C3P(11) C3P(12) C3P(13) C3P(14)
        C3P(15) C3P(16)
#undef C3P
        default: throw Error(-1, "batched clique kernel: no instance for this size");
    }
    HIPCHK(hipGetLastError());
    const int64_t gw = ((sp.nstr + 3) / 4) * ((L + 63) / 64);
    const dim3 g2(xcd_grid(gw), unsigned(nitem));
    const double* const* Ec = E_t;
    if (acc)
        k_clique_gather<double, true, true><<<g2, 256, 0, st>>>(nullptr, nullptr, sp.nstr, L, sp.own.p, sp.nown, 0, L,
                                                               Ls, nullptr, nullptr, sp.nK * sp.M, act, Ec, Y_t);
    else
        k_clique_gather<double, false, true><<<g2, 256, 0, st>>>(nullptr, nullptr, sp.nstr, L, sp.own.p, sp.nown, 0,
                                                                L, Ls, nullptr, nullptr, sp.nK * sp.M, act, Ec, Y_t);
    HIPCHK(hipGetLastError());
}

template <int KS, int TPW, bool SPLIT>
static void batch_ab2_f64(const SpinTables& a, const SpinTables& b, const BatchSigmaTabs& t, hipStream_t st) {
    const size_t sh = size_t(b.nstr) * sizeof(double);  // attribute set by batch_sigma_prepare
    auto kern = k_ab2_f64<KS, TPW, SPLIT>;
    kern<<<dim3(xcd_grid(a.nstr), unsigned(t.nitem)), 512, sh, st>>>(
        nullptr, nullptr, a.nstr, b.nstr, a.into.p, a.ns, a.str.p, b.outT.p, b.npp, b.np, nullptr, (b.np + 15) / 16,
        t.act, t.c, t.s, t.vpk);
    HIPCHK(hipGetLastError());
}

template <int KS, int TPW, bool SPLIT> static void batch_ab2_attr(const SpinTables& b) {
    HIPCHK(hipFuncSetAttribute((const void*)k_ab2_f64<KS, TPW, SPLIT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               int(size_t(b.nstr) * sizeof(double))));
}

void batch_sigma_prepare(const SpinTables& a, const SpinTables& b) {
    const int ks = std::max(1, (a.ns + 15) / 16) * 4;
    const int tpw = ((b.np + 15) / 16 + 3) / 4;
    const bool split = tpw > 1 && (b.np + 15) / 16 == 4 * (tpw - 1) + 1;
#define F64CASE(K, T) \
    if (ks == K && tpw == T) return split ? batch_ab2_attr<K, T, (T > 1)>(b) : batch_ab2_attr<K, T, false>(b);
#define F64ROW(K) F64CASE(K, 1) F64CASE(K, 2) F64CASE(K, 3)
    F64ROW(4) F64ROW(8) F64ROW(12) F64ROW(16) F64ROW(20) F64ROW(24)
#undef F64ROW
#undef F64CASE
    throw Error(-1, "batched alpha-beta kernel: no instance for this shape");
}

static void batch_ab_f64(const SpinTables& a, const SpinTables& b, const BatchSigmaTabs& t, hipStream_t st);

void batch_sigma_f64(const SpinTables& a, const SpinTables& b, int gd_a, int gd_b, const BatchSigmaTabs& t,
                     hipStream_t st) {
    const int64_t Na = a.nstr, Nb = b.nstr;
    const unsigned ni = unsigned(t.nitem);
    // beta same-spin on C^T: C -> C^T, cliques -> Y^T, sigma = (Y^T)^T
    k_transpose64<double><<<dim3(nblk(Nb, 64), nblk(Na, 64), ni), 256, 0, st>>>(nullptr, nullptr, Na, Nb, nullptr,
                                                                               0.0, t.act, t.c, t.ct, nullptr);
    HIPCHK(hipGetLastError());
    batch_clique_f64(b, gd_b, Na, t.act, t.ct, t.e, t.g2b, t.yt, false, t.nitem, st);
    k_transpose64<double><<<dim3(nblk(Na, 64), nblk(Nb, 64), ni), 256, 0, st>>>(nullptr, nullptr, Nb, Na, nullptr,
                                                                               0.0, t.act, t.yt, t.s, nullptr);
    HIPCHK(hipGetLastError());
    // alpha same-spin on C, accumulated into sigma
    batch_clique_f64(a, gd_a, Nb, t.act, t.c, t.e, t.g2a, t.s, true, t.nitem, st);
    batch_ab_f64(a, b, t, st);
}

void batch_sigma_f64_ms0(const SpinTables& a, const SpinTables& b, int gd_a, int eta, const BatchSigmaTabs& t,
                         hipStream_t st) {
    const int64_t N = a.nstr;
    const unsigned ni = unsigned(t.nitem), tg = unsigned((N + 63) / 64);
    // full C of every item from its packed coordinates
    k_ms0_unpack<double, double><<<dim3(tg, tg, ni), 256, 0, st>>>(nullptr, nullptr, N, eta == 1 ? 0 : 1, eta,
                                                                    nullptr, 0.0, t.act, t.c, t.ct);
    HIPCHK(hipGetLastError());
    // S = H_aa C (written) + 1/2 H_ab C (the item tables hold halved integrals)
    BatchSigmaTabs u = t;
    u.c = t.ct;
    u.s = t.yt;
    batch_clique_f64(a, gd_a, N, u.act, u.c, u.e, u.g2a, u.s, false, u.nitem, st);
    batch_ab_f64(a, b, u, st);
    // packed coordinates of sigma = S + eta S^T
    k_ms0_pack<double, double><<<dim3(tg, tg, ni), 256, 0, st>>>(nullptr, nullptr, N, eta == 1 ? 0 : 1, eta,
                                                                  1.41421356237309504880, 2.0, t.act, u.s, t.s);
    HIPCHK(hipGetLastError());
}

static void batch_ab_f64(const SpinTables& a, const SpinTables& b, const BatchSigmaTabs& t, hipStream_t st) {
    const int ks = std::max(1, (a.ns + 15) / 16) * 4;
    const int tpw = ((b.np + 15) / 16 + 3) / 4;
    const bool split = tpw > 1 && (b.np + 15) / 16 == 4 * (tpw - 1) + 1;
#define F64CASE(K, T)                                                                              \
    if (ks == K && tpw == T)                                                                        \
        return split ? batch_ab2_f64<K, T, (T > 1)>(a, b, t, st) : batch_ab2_f64<K, T, false>(a, b, t, st);
#define F64ROW(K) F64CASE(K, 1) F64CASE(K, 2) F64CASE(K, 3)
    F64ROW(4) F64ROW(8) F64ROW(12) F64ROW(16) F64ROW(20) F64ROW(24)
#undef F64ROW
#undef F64CASE
    throw Error(-1, "batched alpha-beta kernel: no instance for this shape");
}

static inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

size_t SigmaEngine::scratch_min_bytes(size_t w) const {
    const SpinTables& a = *a_;
    const SpinTables& b = *b_;
    const size_t nd = size_t(a.nstr) * size_t(b.nstr);
    const size_t vec = b.d > 0 ? 2 * align256(nd * w) : 0;
    size_t e = 0;
    for (int spin = 0; spin < 2; ++spin) {
        const SpinTables& sp = spin ? b : a;
        if (sp.d == 0) continue;
        const bool two = w == 8 ? clique_two_pass<double>(sp, *ints_, spin) : clique_two_pass<float>(sp, *ints_, spin);
        if (!two) continue;
        const int tw = w == 8 ? clique_tile_width<double>(sp) : clique_tile_width<float>(sp);
        e = std::max(e, size_t(sp.nK) * size_t(sp.M) * size_t(tw) * w);
    }
    return vec + e;
}

template <class T>
void SigmaEngine::apply_t(const T* C, T* S, hipStream_t st, unsigned char* ext, size_t ext_bytes) {
    const SpinTables& a = *a_;
    const SpinTables& b = *b_;
    const int64_t Na = a.nstr, Nb = b.nstr, nd = Na * Nb;
    const T ecore = T(ints_->e_core);
    unsigned char* eext = nullptr;  // same-spin intermediate inside the external scratch
    size_t eext_bytes = 0;
    T *ct = nullptr, *yt = nullptr;
    if (ext) {
        if (ext_bytes < scratch_min_bytes(sizeof(T)))
            throw Error(-1, "sigma: this problem needs more scratch than the external region holds");
        size_t off = 0;
        if (b.d > 0) {
            ct = reinterpret_cast<T*>(ext);
            yt = reinterpret_cast<T*>(ext + align256(size_t(nd) * sizeof(T)));
            off = 2 * align256(size_t(nd) * sizeof(T));
        }
        eext = ext + off;
        eext_bytes = ext_bytes - off;
    }
    if (timing_) HIPCHK(hipEventRecord(w_->ev[0], st));
    // beta same-spin on C^T, then sigma = (Y^T) + e_core C   (H0, fci_vector_hamiltonian.cc:84-91)
    if (b.d > 0) {
        if (!ext) {
            DBuf<T>& Ct = scratch<T>(*w_, 0);
            DBuf<T>& Yt = scratch<T>(*w_, 1);
            if (Ct.n < size_t(nd)) Ct.alloc(nd);
            if (Yt.n < size_t(nd)) Yt.alloc(nd);
            ct = Ct.p;
            yt = Yt.p;
        }
        transpose_device<T>(C, ct, Na, Nb, st);
        launch_clique<T>(ct, yt, Na, b, *ints_, 1, false, *w_, st, eext, eext_bytes);
        transpose_add_device<T>(yt, S, Nb, Na, C, ecore, st);
    } else {
        k_scale_copy<T><<<nblk(nd, 256), 256, 0, st>>>(C, S, nd, ecore);
    }
    if (timing_) HIPCHK(hipEventRecord(w_->ev[1], st));
    if (a.d > 0) launch_clique<T>(C, S, Nb, a, *ints_, 0, true, *w_, st, eext, eext_bytes);
    if (timing_) HIPCHK(hipEventRecord(w_->ev[2], st));
    if (a.k > 0 && b.k > 0) launch_ab(C, S, a, b, *ints_, st, false);
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

// Ms = 0 with C^T = eta C: sigma = e_core C + H_aa C + H_bb C + H_ab C where H_bb C = eta (H_aa C)^T
// (the beta same-spin operator is the alpha one on the transposed vector, and the two spins share
// strings and integrals) and H_ab C is eta-symmetric. Built as S = 1/2 e_core C + H_aa C + 1/2 H_ab C
// (packed integrals scaled by 1/2, exact), then S <- S + eta S^T; the eta-symmetric part of H_ab C
// replaces H_ab C itself, which differs from it only by rounding.
template <class T>
void SigmaEngine::apply_sym_t(const T* C, T* S, int eta, hipStream_t st, bool init, bool finish) {
    const SpinTables& a = *a_;
    const SpinTables& b = *b_;
    if (a.nstr != b.nstr || a.k != b.k) throw Error(-1, "sigma: the spin-flip form needs equal alpha and beta strings");
    if (eta != 1 && eta != -1) throw Error(-1, "sigma: eta must be +1 or -1");
    const int64_t N = a.nstr, nd = N * N;
    if (timing_) HIPCHK(hipEventRecord(w_->ev[0], st));
    if (init) {
        k_scale_copy<T><<<nblk(nd, 256), 256, 0, st>>>(C, S, nd, T(0.5 * ints_->e_core));
        HIPCHK(hipGetLastError());
    }
    if (timing_) HIPCHK(hipEventRecord(w_->ev[1], st));
    if (a.d > 0) launch_clique<T>(C, S, N, a, *ints_, 0, true, *w_, st, nullptr, 0);
    if (timing_) HIPCHK(hipEventRecord(w_->ev[2], st));
    if (a.k > 0) launch_ab(C, S, a, b, *ints_, st, true);
    if (finish) sym_add_device<T>(S, N, eta, false, st);
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
void SigmaEngine::apply_sym(const double* C, double* S, int eta, hipStream_t st) {
    apply_sym_t<double>(C, S, eta, st, true, true);
}
void SigmaEngine::apply_sym(const float* C, float* S, int eta, hipStream_t st) {
    apply_sym_t<float>(C, S, eta, st, true, true);
}
void SigmaEngine::apply_sym_half(const double* C, double* S, int eta, hipStream_t st) {
    apply_sym_t<double>(C, S, eta, st, false, false);
}
void SigmaEngine::apply_sym_half(const float* C, float* S, int eta, hipStream_t st) {
    apply_sym_t<float>(C, S, eta, st, false, false);
}

void SigmaEngine::apply(const double* C, double* S, hipStream_t st) { apply_t<double>(C, S, st, nullptr, 0); }
void SigmaEngine::apply(const float* C, float* S, hipStream_t st) { apply_t<float>(C, S, st, nullptr, 0); }
void SigmaEngine::apply_ext(const double* C, double* S, hipStream_t st, void* ext, size_t ext_bytes) {
    if (!ext) throw Error(-1, "sigma: null external scratch");
    apply_t<double>(C, S, st, static_cast<unsigned char*>(ext), ext_bytes);
}
void SigmaEngine::apply_ext(const float* C, float* S, hipStream_t st, void* ext, size_t ext_bytes) {
    if (!ext) throw Error(-1, "sigma: null external scratch");
    apply_t<float>(C, S, st, static_cast<unsigned char*>(ext), ext_bytes);
}

}  // namespace fci
