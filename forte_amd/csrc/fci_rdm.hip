// fci_rdm.hip -- 1-/2-RDMs and <S^2> on gfx950 (FCIVector::compute_rdms, fci_vector_rdm.cc:43-393;
// FCIVector::compute_spin2, fci_vector_spin2.cc:38-75). L and R are N_a x N_b FP64 vectors.
//
// g1[p][q]       = <L| a+p a_q |R>                     one spin, rows of L/R (beta: transposes)
// g2ss[p][q][r][s] = <L| a+p a+q a_s a_r |R>            same spin, through the pair cliques:
//                   sum_K <L|a+p a+q|K><K|a_s a_r|R>, a dense M x M product per hole string K
// g2ab[p][r][q][s] = <L| E^a_pq E^b_rs |R>              per alpha row, singles x beta excitations
// <S^2> = -sum_rs <E^a_sr E^b_rs> + (na-nb)^2/4 + (na+nb)/2
#include "fci_internal.h"

#include <algorithm>
#include <cmath>

namespace fci {

static inline unsigned nblk(int64_t n, int bs) { return unsigned((n + bs - 1) / bs); }

__device__ __forceinline__ void unpack_pair(int P, int& hi, int& lo) {
    int r = 0;
    while ((r + 1) * (r + 2) / 2 <= P) ++r;
    hi = r;
    lo = P - r * (r + 1) / 2;
}

constexpr int RD_BS = 256;
constexpr int RD_MAXE = 192;

// 1-RDM of one spin over rows: block per row J. Partial sums of each single excitation into J
// and of the diagonal, reduced over the block, then added to g[p][q].
__global__ __launch_bounds__(RD_BS) void k_rdm1_rows(const double* __restrict__ L, const double* __restrict__ R,
                                                     int64_t nrow, int64_t ncol, const uint32_t* __restrict__ into,
                                                     int ns, const uint32_t* __restrict__ str, int n,
                                                     double* __restrict__ g) {
    __shared__ uint32_t ent[RD_MAXE];
    __shared__ double red[RD_MAXE + 1][RD_BS / 64];
    const int64_t J = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < ns; i += RD_BS) ent[i] = into[J * ns + i];
    __syncthreads();
    const double* lrow = L + J * ncol;
    for (int e0 = 0; e0 <= ns; e0 += 32) {
        double acc[32];
#pragma unroll
        for (int x = 0; x < 32; ++x) acc[x] = 0.0;
        for (int64_t c = tid; c < ncol; c += RD_BS) {
            double lv = lrow[c];
#pragma unroll
            for (int x = 0; x < 32; ++x) {
                int e = e0 + x;
                if (e < ns) acc[x] += lv * R[int64_t(ent[e] & kAddrMask) * ncol + c];
                else if (e == ns) acc[x] += lv * R[J * ncol + c];
            }
        }
#pragma unroll
        for (int x = 0; x < 32; ++x) {
            double v = acc[x];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
            if (lane == 0 && e0 + x <= ns) red[e0 + x][wave] = v;
        }
    }
    __syncthreads();
    const uint32_t occ = str[J];
    for (int e = tid; e <= ns; e += RD_BS) {
        double s = 0.0;
        for (int w = 0; w < RD_BS / 64; ++w) s += red[e][w];
        if (e < ns) {
            uint32_t t = ent[e];
            int hi, lo;
            unpack_pair(int((t >> 21) & 0x3FFu), hi, lo);
            int p = ((occ >> hi) & 1u) ? hi : lo;  // created orbital is occupied in J
            int q = p == hi ? lo : hi;
            atomicAdd(&g[p * n + q], (t >> 31) ? -s : s);
        } else {
            uint32_t o = occ;
            while (o) {
                int p = __ffs(o) - 1;
                o &= o - 1u;
                atomicAdd(&g[p * n + p], s);
            }
        }
    }
}

// same-spin 2-RDM through pair cliques: acc[m_pair][m'_pair] += s_m s_m' sum_c L[I_m,c] R[I_m',c]
constexpr int RC_TW = 32;
constexpr int RC_MAXM = 96;
// Same-spin 2-RDM through the pair cliques: for clique K, the member-by-member products
// sum_c DL[m][c] DR[m'][c] of the signed member rows. Cliques of more than RC_MAXM members are split
// into RC_MAXM x RC_MAXM member tiles, one block per (K, tile pair): blockIdx.y = ti * nt + tj.
__global__ __launch_bounds__(RD_BS) void k_rdm2_clique(const double* __restrict__ L, const double* __restrict__ R,
                                                       int64_t ncol, const uint32_t* __restrict__ clq, int M,
                                                       double* __restrict__ acc_out, int gdim, int nt) {
    __shared__ uint32_t memI[RC_MAXM], memJ[RC_MAXM];
    __shared__ double DL_[RC_MAXM][RC_TW + 1], DR_[RC_MAXM][RC_TW + 1];
    const int64_t K = blockIdx.x;
    const int ti = blockIdx.y / nt, tj = blockIdx.y % nt;
    const int i0 = ti * RC_MAXM, j0 = tj * RC_MAXM;
    const int Mi = min(RC_MAXM, M - i0), Mj = min(RC_MAXM, M - j0);
    const int tid = threadIdx.x;
    for (int i = tid; i < Mi; i += RD_BS) memI[i] = clq[K * M + i0 + i];
    for (int i = tid; i < Mj; i += RD_BS) memJ[i] = clq[K * M + j0 + i];
    __syncthreads();
    constexpr int PER = (RC_MAXM * RC_MAXM + RD_BS - 1) / RD_BS;
    double acc[PER];
#pragma unroll
    for (int x = 0; x < PER; ++x) acc[x] = 0.0;
    for (int64_t c0 = 0; c0 < ncol; c0 += RC_TW) {
        for (int i = tid; i < max(Mi, Mj) * RC_TW; i += RD_BS) {
            const int m = i / RC_TW, c = i % RC_TW;
            double l = 0.0, r = 0.0;
            if (c0 + c < ncol) {
                if (m < Mi) {
                    const int64_t row = int64_t(memI[m] & kAddrMask);
                    l = L[row * ncol + c0 + c];
                    if (memI[m] >> 31) l = -l;
                }
                if (m < Mj) {
                    const int64_t row = int64_t(memJ[m] & kAddrMask);
                    r = R[row * ncol + c0 + c];
                    if (memJ[m] >> 31) r = -r;
                }
            }
            DL_[m][c] = l;
            DR_[m][c] = r;
        }
        __syncthreads();
#pragma unroll
        for (int x = 0; x < PER; ++x) {
            const int idx = tid + x * RD_BS;
            if (idx < Mi * Mj) {
                const int m = idx / Mj, mp = idx % Mj;
                double sum = 0.0;
                for (int c = 0; c < RC_TW; ++c) sum += DL_[m][c] * DR_[mp][c];
                acc[x] += sum;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int x = 0; x < PER; ++x) {
        const int idx = tid + x * RD_BS;
        if (idx < Mi * Mj) {
            const int m = idx / Mj, mp = idx % Mj;
            const int pm = int((memI[m] >> 21) & 0x3FFu), pmp = int((memJ[mp] >> 21) & 0x3FFu);
            atomicAdd(&acc_out[size_t(pm) * gdim + pmp], acc[x]);
        }
    }
}

// alpha-beta 2-RDM: block per (Ja, chunk of 16 packed beta pairs).
constexpr int RA_PC = 16;
constexpr int RA_TW = 32;
__global__ __launch_bounds__(RD_BS) void k_rdm2_ab(const double* __restrict__ L, const double* __restrict__ R,
                                                   int64_t Na, int64_t Nb, const uint32_t* __restrict__ into_a,
                                                   int ns, const uint32_t* __restrict__ str_a,
                                                   const uint32_t* __restrict__ str_b,
                                                   const uint32_t* __restrict__ out_b, int np, int n,
                                                   double* __restrict__ g2ab) {
    extern __shared__ double sm[];
    __shared__ uint32_t ent[RD_MAXE];
    __shared__ int ep[RD_MAXE], eq[RD_MAXE];
    const int npc = (np + RA_PC - 1) / RA_PC;
    const int64_t Ja = blockIdx.x / npc;
    const int P0 = int(blockIdx.x % npc) * RA_PC;
    const int tid = threadIdx.x;
    const uint32_t occ = str_a[Ja];
    const int ka = __popc(occ);
    const int EF = ns + ka;  // singles, then one diagonal entry per occupied orbital
    double* X = sm;                    // [EF][RA_TW]
    double* Z = sm + EF * RA_TW;       // [2*RA_PC][RA_TW]
    for (int i = tid; i < EF; i += RD_BS) {
        if (i < ns) {
            uint32_t t = into_a[Ja * ns + i];
            ent[i] = t;
            int hi, lo;
            unpack_pair(int((t >> 21) & 0x3FFu), hi, lo);
            int p = ((occ >> hi) & 1u) ? hi : lo;
            ep[i] = p;
            eq[i] = p == hi ? lo : hi;
        } else {
            int j = i - ns, cnt = 0, q = 0;
            for (q = 0; q < n; ++q)
                if ((occ >> q) & 1u) {
                    if (cnt == j) break;
                    ++cnt;
                }
            ent[i] = uint32_t(Ja);
            ep[i] = q;
            eq[i] = q;
        }
    }
    __syncthreads();
    const int NO = EF * 2 * RA_PC;
    constexpr int PER = (RD_MAXE * 2 * RA_PC + RD_BS - 1) / RD_BS;
    double acc[PER];
#pragma unroll
    for (int x = 0; x < PER; ++x) acc[x] = 0.0;
    for (int64_t c0 = 0; c0 < Nb; c0 += RA_TW) {
        for (int i = tid; i < EF * RA_TW; i += RD_BS) {
            int e = i / RA_TW, c = i % RA_TW;
            double v = 0.0;
            if (c0 + c < Nb) {
                uint32_t t = ent[e];
                v = R[int64_t(t & kAddrMask) * Nb + c0 + c];
                if (e < ns && (t >> 31)) v = -v;
            }
            X[e * RA_TW + c] = v;
        }
        for (int i = tid; i < 2 * RA_PC * RA_TW; i += RD_BS) {
            int zr = i / RA_TW, c = i % RA_TW;
            int P = P0 + zr / 2, dir = zr % 2;
            double v = 0.0;
            int64_t Ib = c0 + c;
            if (Ib < Nb && P < np) {
                uint32_t t = out_b[int64_t(P) * Nb + Ib];
                if (t != kInvalid) {
                    int hi, lo;
                    unpack_pair(P, hi, lo);
                    // valid direction annihilates the occupied orbital of the pair
                    int d = (hi == lo) ? 0 : (((str_b[Ib] >> lo) & 1u) ? 0 : 1);
                    if (d == dir) {
                        v = L[Ja * Nb + (t & kAddrMask)];
                        if (t >> 31) v = -v;
                    }
                }
            }
            Z[zr * RA_TW + c] = v;
        }
        __syncthreads();
#pragma unroll
        for (int x = 0; x < PER; ++x) {
            int idx = tid + x * RD_BS;
            if (idx < NO) {
                int e = idx / (2 * RA_PC), zr = idx % (2 * RA_PC);
                double s = 0.0;
                for (int c = 0; c < RA_TW; ++c) s += X[e * RA_TW + c] * Z[zr * RA_TW + c];
                acc[x] += s;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int x = 0; x < PER; ++x) {
        int idx = tid + x * RD_BS;
        if (idx < NO) {
            int e = idx / (2 * RA_PC), zr = idx % (2 * RA_PC);
            int P = P0 + zr / 2, dir = zr % 2;
            if (P >= np) continue;
            int hi, lo;
            unpack_pair(P, hi, lo);
            if (hi == lo && dir == 1) continue;
            int r = dir == 0 ? hi : lo, s = dir == 0 ? lo : hi;  // create r, annihilate s
            int p = ep[e], q = eq[e];
            atomicAdd(&g2ab[((size_t(p) * n + r) * n + q) * n + s], acc[x]);
        }
    }
}

// alpha-beta 2-RDM on FP64 MFMA, block per alpha row Ja: out[e][z] = sum_Ib X[e][Ib] Z[z][Ib] with
// X the rows of R reached from Ja by the alpha single excitations e (signed; one diagonal entry per
// occupied orbital, as above) and Z[2P + dir] the row L[Ja] gathered through the beta excitation
// of pair P in direction dir, for every pair at once: the alpha rows are gathered once per block
// (not once per pair chunk) and each 16 x 16 output tile is a chain of 16x16x4 MFMAs over
// 16-column chunks. LDS holds the chunks k-major with odd row strides, so the 16 lanes reading
// one operand column take consecutive doubles.
// The staging loops issue RM_SU independent loads per thread before they store, and the row L[Ja]
// is staged in LDS when it fits beside the chunks (lrow), so a chunk's gathers do not wait on one
// dependent global load after another.
constexpr int RM_TW = 16;   // columns per chunk (the products' K)
constexpr int RM_SU = 8;    // loads in flight per thread while staging
typedef double rd4 __attribute__((ext_vector_type(4)));
// TPW: output tiles per wave at most; the 16-tile instance fits 256 VGPRs, so two blocks share a CU
// (their chunks and rows of L within half the LDS) and one stages while the other multiplies
template <int TPW>
__global__ __launch_bounds__(RD_BS, TPW <= 16 ? 2 : 1) void k_rdm2_ab_mfma(const double* __restrict__ L, const double* __restrict__ R,
                                                        int64_t Na, int64_t Nb, const uint32_t* __restrict__ into_a,
                                                        int ns, const uint32_t* __restrict__ str_a,
                                                        const uint32_t* __restrict__ str_b,
                                                        const uint32_t* __restrict__ out_b, int np, int n,
                                                        double* __restrict__ g2ab, int lrow) {
    extern __shared__ double sm[];
    __shared__ uint32_t ent[RD_MAXE];
    __shared__ int ep[RD_MAXE], eq[RD_MAXE];
    const int64_t Ja = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t occ = str_a[Ja];
    const int ka = __popc(occ);
    const int EF = ns + ka;
    const int ET = (EF + 15) / 16, ZR = 2 * np, ZT = (ZR + 15) / 16;
    const int XS = ET * 16 + 1, ZS = ZT * 16 + 1;  // k-major rows, odd strides
    double* Xs = sm;                 // [RM_TW][XS]
    double* Zs = sm + RM_TW * XS;    // [RM_TW][ZS]
    double* Ls = Zs + RM_TW * ZS;    // [Nb] when lrow
    const double* Lrow = L + Ja * Nb;
    if (lrow) {
        for (int64_t i = tid; i < Nb; i += RD_BS) Ls[i] = Lrow[i];
        Lrow = Ls;
    }
    for (int i = tid; i < EF; i += RD_BS) {
        if (i < ns) {
            const uint32_t t = into_a[Ja * ns + i];
            ent[i] = t;
            int hi, lo;
            unpack_pair(int((t >> 21) & 0x3FFu), hi, lo);
            const int p = ((occ >> hi) & 1u) ? hi : lo;
            ep[i] = p;
            eq[i] = p == hi ? lo : hi;
        } else {
            int j = i - ns, cnt = 0, q = 0;
            for (q = 0; q < n; ++q)
                if ((occ >> q) & 1u) {
                    if (cnt == j) break;
                    ++cnt;
                }
            ent[i] = uint32_t(Ja);
            ep[i] = q;
            eq[i] = q;
        }
    }
    // padding rows of the operands stay zero for the whole block
    for (int i = tid; i < RM_TW * XS; i += RD_BS) Xs[i] = 0.0;
    for (int i = tid; i < RM_TW * ZS; i += RD_BS) Zs[i] = 0.0;
    __syncthreads();
    const int ntile = ET * ZT;
    rd4 acc[TPW];
#pragma unroll
    for (int t = 0; t < TPW; ++t) acc[t] = rd4{0.0, 0.0, 0.0, 0.0};
    for (int64_t c0 = 0; c0 < Nb; c0 += RM_TW) {
        for (int i0 = tid; i0 < EF * RM_TW; i0 += RM_SU * RD_BS) {
            double v[RM_SU];
#pragma unroll
            for (int u = 0; u < RM_SU; ++u) {
                const int i = i0 + u * RD_BS, e = i / RM_TW, c = i % RM_TW;
                v[u] = 0.0;
                if (i < EF * RM_TW && c0 + c < Nb) v[u] = R[int64_t(ent[e] & kAddrMask) * Nb + c0 + c];
            }
#pragma unroll
            for (int u = 0; u < RM_SU; ++u) {
                const int i = i0 + u * RD_BS, e = i / RM_TW, c = i % RM_TW;
                if (i < EF * RM_TW) Xs[c * XS + e] = (e < ns && (ent[e] >> 31)) ? -v[u] : v[u];
            }
        }
        for (int i0 = tid; i0 < ZR * RM_TW; i0 += RM_SU * RD_BS) {
            uint32_t t[RM_SU];
#pragma unroll
            for (int u = 0; u < RM_SU; ++u) {
                const int i = i0 + u * RD_BS, zr = i / RM_TW, c = i % RM_TW;
                t[u] = kInvalid;
                if (i < ZR * RM_TW && c0 + c < Nb) {
                    const int P = zr / 2, dir = zr % 2;
                    const int64_t Ib = c0 + c;
                    int hi, lo;
                    unpack_pair(P, hi, lo);
                    // valid direction annihilates the occupied orbital of the pair
                    const int d = (hi == lo) ? 0 : (((str_b[Ib] >> lo) & 1u) ? 0 : 1);
                    if (d == dir) t[u] = out_b[int64_t(P) * Nb + Ib];
                }
            }
            double v[RM_SU];
#pragma unroll
            for (int u = 0; u < RM_SU; ++u) v[u] = t[u] != kInvalid ? Lrow[t[u] & kAddrMask] : 0.0;
#pragma unroll
            for (int u = 0; u < RM_SU; ++u) {
                const int i = i0 + u * RD_BS, zr = i / RM_TW, c = i % RM_TW;
                if (i < ZR * RM_TW) Zs[c * ZS + zr] = (t[u] != kInvalid && (t[u] >> 31)) ? -v[u] : v[u];
            }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < RM_TW / 4; ++ks) {
            const int k = ks * 4 + (lane >> 4);
#pragma unroll
            for (int t = 0; t < TPW; ++t) {
                const int tile = wave + 4 * t;
                if (tile < ntile) {
                    const int et = tile / ZT, zt = tile - et * ZT;
                    const double a = Xs[k * XS + et * 16 + (lane & 15)];
                    const double b = Zs[k * ZS + zt * 16 + (lane & 15)];
                    acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[t], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    // D[row 4 v + lane / 16][column lane % 16] of each tile: row = excitation e, column = z
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
        const int tile = wave + 4 * t;
        if (tile >= ntile) continue;
        const int et = tile / ZT, zt = tile - et * ZT;
        const int zr = zt * 16 + (lane & 15);
        if (zr >= ZR) continue;
        const int P = zr / 2, dir = zr % 2;
        int hi, lo;
        unpack_pair(P, hi, lo);
        if (hi == lo && dir == 1) continue;
        const int r = dir == 0 ? hi : lo, s = dir == 0 ? lo : hi;  // create r, annihilate s
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int e = et * 16 + 4 * v + (lane >> 4);
            if (e < EF) atomicAdd(&g2ab[((size_t(ep[e]) * n + r) * n + eq[e]) * n + s], acc[t][v]);
        }
    }
}

// <S^2> partial sums, block per alpha row Ja
__global__ __launch_bounds__(RD_BS) void k_spin2(const double* __restrict__ L, const double* __restrict__ R,
                                                 int64_t Na, int64_t Nb, const uint32_t* __restrict__ into_a, int ns,
                                                 const uint32_t* __restrict__ str_a, const uint32_t* __restrict__ str_b,
                                                 const uint32_t* __restrict__ out_b, double* __restrict__ part) {
    __shared__ uint32_t ent[RD_MAXE];
    __shared__ int ep[RD_MAXE];
    __shared__ double red[RD_BS / 64];
    const int64_t Ja = blockIdx.x;
    const int tid = threadIdx.x;
    const uint32_t occ = str_a[Ja];
    for (int i = tid; i < ns; i += RD_BS) {
        uint32_t t = into_a[Ja * ns + i];
        ent[i] = t;
        int hi, lo;
        unpack_pair(int((t >> 21) & 0x3FFu), hi, lo);
        ep[i] = ((occ >> hi) & 1u) ? hi : lo;
    }
    __syncthreads();
    double s = 0.0;
    for (int64_t Ib = tid; Ib < Nb; Ib += RD_BS) {
        const uint32_t ob = str_b[Ib];
        // diagonal: n^a_r n^b_r
        double d = L[Ja * Nb + Ib] * R[Ja * Nb + Ib];
        s += d * double(__popc(occ & ob));
        for (int e = 0; e < ns; ++e) {
            uint32_t t = ent[e];
            int p = ep[e];
            if (!((ob >> p) & 1u)) continue;  // beta must annihilate p
            uint32_t o = out_b[int64_t((t >> 21) & 0x3FFu) * Nb + Ib];
            if (o == kInvalid) continue;
            double v = L[Ja * Nb + (o & kAddrMask)] * R[int64_t(t & kAddrMask) * Nb + Ib];
            if ((t >> 31) ^ (o >> 31)) v = -v;
            s += v;
        }
    }
    const int lane = tid & 63, wave = tid >> 6;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) red[wave] = s;
    __syncthreads();
    if (tid == 0) {
        double t = 0.0;
        for (int w = 0; w < RD_BS / 64; ++w) t += red[w];
        part[Ja] = t;
    }
}

// ---- 3-RDMs through hole strings ----------------------------------------------------------------
// D[x][y] = sum_c XL[x][c] XR[y][c] per hole string K, accumulated into a compact output.
// MODE 0 (three electrons of one spin): x = member triple T of K (rows I = K + T with the sign of
//   a+T0 a+T1 a+T2 |K>), c = column (the other spin's string), out index = rank of T.
// MODE 1 (two electrons of the first spin, one of the second): x = (member pair P of K, orbital r),
//   c = (k2-1)-electron string Lb of the second spin, XL[x][c] = s_P s_{Lb,r} L[K + P][Lb + r]
//   (zero when r is in Lb), out index = rank(P) * n + r.
constexpr int R3_T = 64, R3_C = 32;
template <int MODE>
__global__ __launch_bounds__(256) void k_rdm3_hole(const double* __restrict__ L, const double* __restrict__ R,
                                                   int64_t ncol, const uint32_t* __restrict__ mem,
                                                   const uint32_t* __restrict__ midx, int M,
                                                   const uint32_t* __restrict__ tab1, int64_t nLb, int n,
                                                   int nxt, double* __restrict__ out, int64_t ldo) {
    __shared__ double XL[R3_T][R3_C + 1], XR[R3_T][R3_C + 1];
    const int64_t g = blockIdx.x;
    const int xt = blockIdx.y % nxt, yt = blockIdx.y / nxt;
    const int X = MODE == 0 ? M : M * n;
    const int64_t nc = MODE == 0 ? ncol : nLb;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const uint32_t* gm = mem + g * M;
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
    // the element (x, c) of the gathered operand V (L or R)
    auto elem = [&](const double* V, int x, int64_t c) -> double {
        if (x >= X || c >= nc) return 0.0;
        const int m = MODE == 0 ? x : x / n;
        const uint32_t e = gm[m];
        const double sm = (e >> 31) ? -1.0 : 1.0;
        const int64_t row = int64_t(e & 0x7FFFFFFFu);
        if (MODE == 0) return sm * V[row * ncol + c];
        const uint32_t t = tab1[c * n + (x % n)];
        if (t == kInvalid) return 0.0;
        return ((t >> 31) ? -sm : sm) * V[row * ncol + int64_t(t & 0x7FFFFFFFu)];
    };
    for (int64_t c0 = 0; c0 < nc; c0 += R3_C) {
        for (int i = tid; i < R3_T * R3_C; i += 256) {
            const int xr = i / R3_C, cc = i % R3_C;
            XL[xr][cc] = elem(L, xt * R3_T + xr, c0 + cc);
            XR[xr][cc] = elem(R, yt * R3_T + xr, c0 + cc);
        }
        __syncthreads();
#pragma unroll 4
        for (int cc = 0; cc < R3_C; ++cc) {
            double a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = XL[tx * 4 + i][cc];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = XR[ty * 4 + j][cc];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int x = xt * R3_T + tx * 4 + i;
        if (x >= X) continue;
        const int64_t ox = MODE == 0 ? midx[g * M + x] : int64_t(midx[g * M + x / n]) * n + x % n;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int y = yt * R3_T + ty * 4 + j;
            if (y >= X) continue;
            const int64_t oy = MODE == 0 ? midx[g * M + y] : int64_t(midx[g * M + y / n]) * n + y % n;
            atomicAdd(&out[ox * ldo + oy], acc[i][j]);
        }
    }
}

namespace {
// (-1)^{#bits of s below orbital p}
inline int sign_below_h(uint32_t s, int p) { return (__builtin_popcount(s & ((1u << p) - 1u)) & 1) ? -1 : 1; }
inline int64_t rank_h(uint32_t s, int n) {
    int64_t r = 0;
    int t = 0;
    for (int i = n - 1; i >= 0; --i)
        if ((s >> i) & 1u) {
            ++t;
            r += int64_t(binom_host(n - 1 - i, t));
        }
    return r;
}
// all k-subsets of n orbitals as bit strings
std::vector<uint32_t> subsets(int n, int k) {
    std::vector<uint32_t> out;
    if (k < 0 || k > n) return out;
    for (uint32_t s = 0; s < (1u << n); ++s)
        if (__builtin_popcount(s) == k) out.push_back(s);
    return out;
}
// hole-string members: for each (k - d)-string K and each d-subset D of vir(K): entry
// addr(K + D) | (sign of a+D0 ... a+D(d-1) |K> < 0) << 31, and the colex rank of D (0-based)
void hole_members(int n, int k, int d, std::vector<uint32_t>& mem, std::vector<uint32_t>& idx, int& M, int64_t& nK) {
    mem.clear();
    idx.clear();
    auto Ks = subsets(n, k - d);
    nK = int64_t(Ks.size());
    M = int(binom_host(n - (k - d), d));
    for (uint32_t K : Ks) {
        int cnt = 0;
        for (uint32_t D : subsets(n, d)) {
            if (D & K) continue;
            int sg = 1;
            int64_t rk = 0;
            int t = 0;
            for (int p = 0; p < n; ++p)
                if ((D >> p) & 1u) {
                    sg *= sign_below_h(K, p);
                    ++t;
                    rk += int64_t(binom_host(p, t));
                }
            mem.push_back(uint32_t(rank_h(K | D, n)) | (sg < 0 ? 0x80000000u : 0u));
            idx.push_back(uint32_t(rk));
            ++cnt;
        }
        if (cnt != M) throw Error(-1, "rdm3: member count mismatch");
    }
}
}  // namespace

// Compact 3-RDM blocks of one orientation (rows of L/R are strings of the first spin, na1
// electrons; columns strings of the second spin, na2 electrons):
//   c3 (C(n,3)^2): <L| a+ a+ a+ a a a |R> of the first spin, sorted triples;
//   c21 ((C(n,2) n)^2): <L| a+P a+P' a+r(2) a_u(2) a_S' a_S |R>, sorted pairs of the first spin.
void rdm3_compact_device(int n, int na1, int na2, const double* L, const double* R, int64_t nrow, int64_t ncol,
                         double* c3, double* c21, hipStream_t st) {
    const int64_t n3 = int64_t(binom_host(n, 3)), n2p = int64_t(binom_host(n, 2)) * n;
    if (c3) HIPCHK(hipMemsetAsync(c3, 0, size_t(n3 * n3) * sizeof(double), st));
    if (c21) HIPCHK(hipMemsetAsync(c21, 0, size_t(n2p * n2p) * sizeof(double), st));
    std::vector<uint32_t> mem, idx;
    int M = 0;
    int64_t nK = 0;
    if (c3 && na1 >= 3) {
        hole_members(n, na1, 3, mem, idx, M, nK);
        DBuf<uint32_t> dm(mem.size()), di(idx.size());
        dm.upload(mem.data(), mem.size());
        di.upload(idx.data(), idx.size());
        const int nxt = (M + R3_T - 1) / R3_T;
        k_rdm3_hole<0><<<dim3(unsigned(nK), unsigned(nxt * nxt)), 256, 0, st>>>(L, R, ncol, dm.p, di.p, M, nullptr, 0,
                                                                               n, nxt, c3, n3);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
    }
    if (c21 && na1 >= 2 && na2 >= 1) {
        hole_members(n, na1, 2, mem, idx, M, nK);
        DBuf<uint32_t> dm(mem.size()), di(idx.size());
        dm.upload(mem.data(), mem.size());
        di.upload(idx.data(), idx.size());
        // second spin: (na2 - 1)-strings Lb and orbital r -> addr(Lb + r) | sign << 31
        auto Lbs = subsets(n, na2 - 1);
        std::vector<uint32_t> t1(Lbs.size() * size_t(n), kInvalid);
        for (size_t i = 0; i < Lbs.size(); ++i)
            for (int r = 0; r < n; ++r)
                if (!((Lbs[i] >> r) & 1u))
                    t1[i * n + r] = uint32_t(rank_h(Lbs[i] | (1u << r), n)) | (sign_below_h(Lbs[i], r) < 0 ? 0x80000000u : 0u);
        DBuf<uint32_t> dt(t1.size());
        dt.upload(t1.data(), t1.size());
        const int nxt = (M * n + R3_T - 1) / R3_T;
        k_rdm3_hole<1><<<dim3(unsigned(nK), unsigned(nxt * nxt)), 256, 0, st>>>(
            L, R, ncol, dm.p, di.p, M, dt.p, int64_t(Lbs.size()), n, nxt, c21, n2p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
    }
    (void)nrow;
}

// ---- launchers -------------------------------------------------------------------------------
void rdm1_device(const SpinTables& a, const SpinTables& b, const double* L, const double* R, double* g1a,
                 double* g1b, hipStream_t st) {
    const int n = a.n;
    const int64_t Na = a.nstr, Nb = b.nstr;
    HIPCHK(hipMemsetAsync(g1a, 0, size_t(n) * n * sizeof(double), st));
    HIPCHK(hipMemsetAsync(g1b, 0, size_t(n) * n * sizeof(double), st));
    if (a.ns + 1 > RD_MAXE || b.ns + 1 > RD_MAXE) throw Error(-1, "rdm1: too many excitations per string");
    if (a.k > 0) k_rdm1_rows<<<unsigned(Na), RD_BS, 0, st>>>(L, R, Na, Nb, a.into.p, a.ns, a.str.p, n, g1a);
    if (b.k > 0) {
        DBuf<double> Lt(size_t(Na) * Nb), Rt(size_t(Na) * Nb);
        transpose_device<double>(L, Lt.p, Na, Nb, st);
        transpose_device<double>(R, Rt.p, Na, Nb, st);
        k_rdm1_rows<<<unsigned(Nb), RD_BS, 0, st>>>(Lt.p, Rt.p, Nb, Na, b.into.p, b.ns, b.str.p, n, g1b);
        HIPCHK(hipStreamSynchronize(st));
    }
    HIPCHK(hipGetLastError());
}

void rdm2_same_device(const SpinTables& a, const SpinTables& b, const double* L, const double* R, double* acc_a,
                      double* acc_b, hipStream_t st) {
    // acc_a / acc_b: npair x npair accumulators (pair index p(p-1)/2+q for p>q)
    const int64_t Na = a.nstr, Nb = b.nstr;
    const int npair = a.n * (a.n - 1) / 2;
    HIPCHK(hipMemsetAsync(acc_a, 0, size_t(npair) * npair * sizeof(double), st));
    HIPCHK(hipMemsetAsync(acc_b, 0, size_t(npair) * npair * sizeof(double), st));
    if (a.d == 2) {
        const int nta = (a.M + RC_MAXM - 1) / RC_MAXM;
        k_rdm2_clique<<<dim3(unsigned(a.nK), unsigned(nta * nta)), RD_BS, 0, st>>>(L, R, Nb, a.clq.p, a.M, acc_a, npair,
                                                                                  nta);
    }
    if (b.d == 2) {
        const int ntb = (b.M + RC_MAXM - 1) / RC_MAXM;
        DBuf<double> Lt(size_t(Na) * Nb), Rt(size_t(Na) * Nb);
        transpose_device<double>(L, Lt.p, Na, Nb, st);
        transpose_device<double>(R, Rt.p, Na, Nb, st);
        k_rdm2_clique<<<dim3(unsigned(b.nK), unsigned(ntb * ntb)), RD_BS, 0, st>>>(Lt.p, Rt.p, Na, b.clq.p, b.M, acc_b,
                                                                                   npair, ntb);
        HIPCHK(hipStreamSynchronize(st));
    }
    HIPCHK(hipGetLastError());
}

void rdm2_ab_device(const SpinTables& a, const SpinTables& b, const double* L, const double* R, double* g2ab,
                    hipStream_t st) {
    const int n = a.n;
    HIPCHK(hipMemsetAsync(g2ab, 0, size_t(n) * n * n * n * sizeof(double), st));
    if (a.k == 0 || b.k == 0) return;
    if (a.ns + a.k > RD_MAXE) throw Error(-1, "rdm2ab: too many excitations per string");
    // MFMA kernel when every output tile of a row has a wave slot and the chunks fit in LDS
    const int ET = (a.ns + a.k + 15) / 16, ZT = (2 * b.np + 15) / 16;
    const size_t shm = size_t(RM_TW) * (ET * 16 + 1 + ZT * 16 + 1) * sizeof(double);
    const size_t shl = shm + size_t(b.nstr) * sizeof(double);  // with the row L[Ja] staged
    if (ET * ZT <= 4 * 24 && shm <= 150 * 1024) {  // 24: tiles per wave of the largest instance
        const int lrow = shl <= 150 * 1024;
        const size_t sh2 = lrow ? shl : shm;
        auto run = [&](auto kern) {
            HIPCHK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, int(sh2)));
            kern<<<unsigned(a.nstr), RD_BS, sh2, st>>>(L, R, a.nstr, b.nstr, a.into.p, a.ns, a.str.p, b.str.p, b.out.p,
                                                       b.np, n, g2ab, lrow);
        };
        if (ET * ZT <= 4 * 16) run(k_rdm2_ab_mfma<16>);
        else run(k_rdm2_ab_mfma<24>);
        HIPCHK(hipGetLastError());
        return;
    }
    const int npc = (b.np + RA_PC - 1) / RA_PC;
    size_t sh = size_t(a.ns + a.k + 2 * RA_PC) * RA_TW * sizeof(double);
    k_rdm2_ab<<<unsigned(a.nstr * npc), RD_BS, sh, st>>>(L, R, a.nstr, b.nstr, a.into.p, a.ns, a.str.p, b.str.p,
                                                        b.out.p, b.np, n, g2ab);
    HIPCHK(hipGetLastError());
}

double spin2_device(const SpinTables& a, const SpinTables& b, const double* L, const double* R, hipStream_t st) {
    DBuf<double> part(a.nstr);
    if (a.ns + 1 > RD_MAXE) throw Error(-1, "spin2: too many excitations per string");
    k_spin2<<<unsigned(a.nstr), RD_BS, 0, st>>>(L, R, a.nstr, b.nstr, a.into.p, a.ns, a.str.p, b.str.p, b.out.p,
                                               part.p);
    HIPCHK(hipGetLastError());
    std::vector<double> h = part.to_host();
    double s = 0.0;
    for (double x : h) s += x;
    double na = a.k, nb = b.k;
    return -s + 0.25 * (na - nb) * (na - nb) + 0.5 * (na + nb);
}

}  // namespace fci
