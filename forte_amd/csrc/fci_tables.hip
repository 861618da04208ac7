// fci_tables.hip -- strings, excitation tables, H diagonal and k-smallest selection on gfx950.
//
// String order. Forte enumerates k-of-n strings with std::next_permutation over bits 0..n-1
// starting from 0..01..1 (forte/fci/fci_string_lists.cc:258-293) and addresses them by order of
// appearance (fci_string_address.cc:45-50). That order is the colex rank of the bit-reversed
// occupation: with c_1 < ... < c_k the values n-1-i over set bits i, address = sum_t C(c_t, t).
// Every table below is therefore built in parallel: a thread unranks its index into a bit
// pattern, or ranks a bit pattern into an address, with a small binomial table.
//
// Forte-order lists (parity only). make_vo / make_oo / make_vvoo
// (fci_string_vo_list.cc:90-140, fci_string_oo_list.cc:72-115, fci_string_vvoo_list.cc:88-187)
// enumerate the remaining orbitals with the same next_permutation order, so entry t of a list
// is "unrank t in the reduced space, then re-insert the fixed orbitals".
#include "fci_internal.h"

#include <algorithm>
#include <cmath>
#include <functional>

namespace fci {

[[noreturn]] void hip_fail(hipError_t e, const char* what, const char* file, int line) {
    char buf[512];
    snprintf(buf, sizeof(buf), "HIP error %d (%s) in %s at %s:%d", int(e), hipGetErrorString(e),
             what, file, line);
    throw Error(-3, buf);
}

std::vector<uint32_t> binom_table() {
    std::vector<uint32_t> b(33 * 33, 0);
    for (int a = 0; a <= 32; ++a)
        for (int c = 0; c <= a; ++c) b[a * 33 + c] = uint32_t(binom_host(a, c));
    return b;
}

__device__ __forceinline__ uint32_t d_rank(uint32_t s, int n, const uint32_t* __restrict__ B) {
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

__device__ __forceinline__ uint32_t d_unrank(uint32_t a, int n, int k, const uint32_t* __restrict__ B) {
    uint32_t s = 0;
    int c = n;
    for (int t = k; t >= 1; --t) {
        int cc = c - 1;
        while (B[cc * 33 + t] > a) --cc;
        a -= B[cc * 33 + t];
        s |= 1u << (n - 1 - cc);
        c = cc;
    }
    return s;
}

// (-1)^(set bits of s strictly between positions x and y)
__device__ __forceinline__ uint32_t d_between_parity(uint32_t s, int x, int y) {
    int lo = x < y ? x : y, hi = x < y ? y : x;
    if (hi - lo < 2) return 0u;
    uint32_t mask = ((1u << hi) - 1u) & ~((2u << lo) - 1u);
    return __popc(s & mask) & 1u;
}

__device__ __forceinline__ int d_pair_packed(int p, int q) {
    int hi = p > q ? p : q, lo = p > q ? q : p;
    return hi * (hi + 1) / 2 + lo;
}

__global__ void k_strings(uint32_t* out, int64_t nstr, int n, int k, const uint32_t* B) {
    int64_t a = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (a < nstr) out[a] = d_unrank(uint32_t(a), n, k, B);
}

// singles into J (p occupied in J, q empty in J): J = sign a+_p a_q I
__global__ void k_into(const uint32_t* __restrict__ str, uint32_t* out, int64_t nstr, int n, int k,
                       int ns, const uint32_t* B) {
    int64_t J = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (J >= nstr) return;
    uint32_t s = str[J];
    uint32_t* o = out + J * ns;
    int e = 0;
    for (int p = 0; p < n; ++p) {
        if (!((s >> p) & 1u)) continue;
        for (int q = 0; q < n; ++q) {
            if ((s >> q) & 1u) continue;
            uint32_t I = s ^ (1u << p) ^ (1u << q);
            uint32_t neg = d_between_parity(s, p, q);
            o[e++] = d_rank(I, n, B) | (uint32_t(d_pair_packed(p, q)) << 21) | (neg << 31);
        }
    }
}

// out of I along the packed pair P = {r >= s}; stored pair-major out[P*nstr + I]
__global__ void k_out(const uint32_t* __restrict__ str, uint32_t* out, int64_t nstr, int n, int np,
                      const uint32_t* B) {
    int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nstr * np) return;
    int64_t I = t / np;
    int P = int(t - I * np);
    int r = 0;
    while ((r + 1) * (r + 2) / 2 <= P) ++r;
    int s = P - r * (r + 1) / 2;
    uint32_t x = str[I];
    uint32_t br = (x >> r) & 1u, bs = (x >> s) & 1u;
    uint32_t v = kInvalid;
    if (r == s) {
        if (br) v = uint32_t(I);
    } else if (br != bs) {
        uint32_t Jb = x ^ (1u << r) ^ (1u << s);
        v = d_rank(Jb, n, B) | (d_between_parity(x, r, s) << 31);
    }
    out[int64_t(P) * nstr + I] = v;  // pair-major: consecutive strings are adjacent
}

// pair cliques: hole string K (k-d electrons), members K + d electrons of vir(K)
__global__ void k_clique(uint32_t* out, int64_t nK, int n, int k, int d, int M, const uint32_t* B) {
    int64_t x = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (x >= nK) return;
    uint32_t K = d_unrank(uint32_t(x), n, k - d, B);
    uint32_t* o = out + x * M;
    int m = 0;
    if (d == 2) {
        for (int a = 0; a < n; ++a) {
            if ((K >> a) & 1u) continue;
            for (int b = 0; b < a; ++b) {
                if ((K >> b) & 1u) continue;
                uint32_t I = K | (1u << a) | (1u << b);
                // a+_a a+_b |K>: (-1)^(n_<b(K)) (-1)^(n_<a(K)+1)
                uint32_t nb_ = __popc(K & ((1u << b) - 1u)), na_ = __popc(K & ((1u << a) - 1u));
                uint32_t neg = (na_ + nb_ + 1u) & 1u;
                o[m++] = d_rank(I, n, B) | (uint32_t(a * (a - 1) / 2 + b) << 21) | (neg << 31);
            }
        }
    } else {
        for (int a = 0; a < n; ++a) {
            if ((K >> a) & 1u) continue;
            uint32_t I = K | (1u << a);
            uint32_t neg = __popc(K & ((1u << a) - 1u)) & 1u;
            o[m++] = d_rank(I, n, B) | (uint32_t(a) << 21) | (neg << 31);
        }
    }
}

// owners: own[I][j] = K * M + m for the C(k,d) hole strings K of string I (the d-subsets of
// occ(I), pairs a > b in ascending (a, b) order) and the member index m of I in clique K.
__global__ void k_clique_owner(const uint32_t* __restrict__ str, int64_t nstr, int n, int k, int d, int M,
                               const uint32_t* __restrict__ B, uint32_t* __restrict__ own, int nown) {
    int64_t I = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (I >= nstr) return;
    const uint32_t s = str[I];
    uint32_t* o = own + I * nown;
    int j = 0;
    for (int a = 0; a < n; ++a) {
        if (!((s >> a) & 1u)) continue;
        if (d == 1) {
            o[j++] = uint32_t(a);  // K is the empty string (rank 0); member index = orbital
            continue;
        }
        for (int b = 0; b < a; ++b) {
            if (!((s >> b) & 1u)) continue;
            const uint32_t K = s & ~((1u << a) | (1u << b));
            // position of a and b among the unoccupied orbitals of K
            const int pa = a - __popc(K & ((1u << a) - 1u));
            const int pb = b - __popc(K & ((1u << b) - 1u));
            o[j++] = d_rank(K, n, B) * uint32_t(M) + uint32_t(pa * (pa - 1) / 2 + pb);
        }
    }
    (void)k;
}

static inline unsigned nblk(int64_t n, int bs) { return unsigned((n + bs - 1) / bs); }

// o16 in the paired-column MFMA tile layout used by the alpha-beta kernels: strings in groups of
// 32, where lane (g, c) of a wave handles strings 32 it + 2c and 32 it + 2c + 1 and pairs
// 16 tile + 4 g + v (v = 0..3):
//   element (((it * ntile + tile) * 64 + 16 g + c) * 2 + p) * 4 + v,  string 32 it + 2c + p,
// so the 8 targets a lane needs for one tile are one 16-byte word and a wave's 64 words are
// contiguous. Strings beyond nstr and pairs beyond np hold 0x7FFF (none).
__global__ void k_out_t16(const uint32_t* __restrict__ out, int64_t nstr, int np, int npp, uint16_t* __restrict__ o16) {
    const int ntile = npp / 16;
    const int64_t nrow = (nstr + 31) / 32 * 32;
    const int64_t x = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (x >= nrow * npp) return;
    const int v = int(x % 4);
    const int pp = int((x / 4) % 2);
    const int lane = int((x / 8) % 64);
    const int64_t q = x / 512;  // it * ntile + tile
    const int tile = int(q % ntile);
    const int64_t it = q / ntile;
    const int64_t I = it * 32 + 2 * (lane % 16) + pp;
    const int P = tile * 16 + 4 * (lane / 16) + v;
    uint16_t val = 0x7FFFu;
    if (P < np && I < nstr) {
        const uint32_t t = out[int64_t(P) * nstr + I];
        if (t != kInvalid) val = uint16_t((t & kAddrMask) | ((t >> 31) << 15));
    }
    o16[x] = val;
}

void build_spin_tables(SpinTables& t, int n, int k, const uint32_t* B, hipStream_t st) {
    t.n = n;
    t.k = k;
    t.nstr = int64_t(binom_host(n, k));
    if (t.nstr > int64_t(kAddrMask)) throw Error(-1, "string space too large for 21-bit addresses");
    t.str.alloc(t.nstr);
    k_strings<<<nblk(t.nstr, 256), 256, 0, st>>>(t.str.p, t.nstr, n, k, B);
    t.ns = k * (n - k);
    if (t.ns > 0) {
        t.into.alloc(size_t(t.nstr) * t.ns);
        k_into<<<nblk(t.nstr, 128), 128, 0, st>>>(t.str.p, t.into.p, t.nstr, n, k, t.ns, B);
    }
    t.np = n * (n + 1) / 2;
    t.out.alloc(size_t(t.nstr) * t.np);
    k_out<<<nblk(t.nstr * t.np, 256), 256, 0, st>>>(t.str.p, t.out.p, t.nstr, n, t.np, B);
    if (t.nstr <= 32767) {
        t.npp = (t.np + 15) / 16 * 16;
        const int64_t nrow = (t.nstr + 31) / 32 * 32;
        t.outT.alloc(size_t(nrow) * t.npp);
        k_out_t16<<<nblk(nrow * t.npp, 256), 256, 0, st>>>(t.out.p, t.nstr, t.np, t.npp, t.outT.p);
    }
    t.d = k >= 2 ? 2 : (k == 1 ? 1 : 0);
    if (t.d > 0) {
        t.nK = int64_t(binom_host(n, k - t.d));
        t.M = int(binom_host(n - k + t.d, t.d));
        t.clq.alloc(size_t(t.nK) * t.M);
        k_clique<<<nblk(t.nK, 64), 64, 0, st>>>(t.clq.p, t.nK, n, k, t.d, t.M, B);
        t.nown = int(binom_host(k, t.d));
        if (uint64_t(t.nK) * uint64_t(t.M) > 0xFFFFFFFFull) throw Error(-1, "clique table too large");
        t.own.alloc(size_t(t.nstr) * t.nown);
        k_clique_owner<<<nblk(t.nstr, 128), 128, 0, st>>>(t.str.p, t.nstr, n, k, t.d, t.M, B, t.own.p, t.nown);
    }
    HIPCHK(hipGetLastError());
}

// ---- point-group classes of the strings (FCIStringLists::make_fci_strings with symmetry) ------
// Forte enumerates the strings in next_permutation order and files each one under its irrep, its
// address being its position among the strings of that irrep (fci_string_lists.cc:258-293,
// fci_string_address.cc:36-59). The C1 order is the same enumeration, so the address within an
// irrep is the number of same-irrep strings before it: per block, a ballot per irrep and wave
// gives each string its rank in the wave and the wave totals; block totals are prefix-summed
// across blocks on the host (one-time table work).
constexpr int SC_BS = 256, SC_NIRREP = 8;
__global__ __launch_bounds__(SC_BS) void k_string_irrep(const uint32_t* __restrict__ str, int64_t nstr,
                                                        const uint32_t* __restrict__ orbsym_bits,
                                                        uint8_t* __restrict__ irrep, uint32_t* __restrict__ rank,
                                                        uint32_t* __restrict__ block_count) {
    __shared__ uint32_t wcnt[SC_BS / 64][SC_NIRREP];
    const int64_t I = int64_t(blockIdx.x) * SC_BS + threadIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int h = -1;
    if (I < nstr) {
        // irrep = XOR of the occupied orbitals' irreps; orbsym_bits[b] holds bit b of every orbital's irrep
        const uint32_t s = str[I];
        h = (__builtin_popcount(s & orbsym_bits[0]) & 1) | ((__builtin_popcount(s & orbsym_bits[1]) & 1) << 1) |
            ((__builtin_popcount(s & orbsym_bits[2]) & 1) << 2);
        irrep[I] = uint8_t(h);
    }
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    uint32_t in_wave = 0;
    for (int g = 0; g < SC_NIRREP; ++g) {
        const uint64_t m = __ballot(h == g);
        if (h == g) in_wave = uint32_t(__popcll(m & below));
        if (lane == 0) wcnt[w][g] = uint32_t(__popcll(m));
    }
    __syncthreads();
    if (I < nstr) {
        uint32_t off = 0;
        for (int v = 0; v < w; ++v) off += wcnt[v][h];
        rank[I] = off + in_wave;  // within the block; the block offset is added afterwards
    }
    if (threadIdx.x < SC_NIRREP) {
        uint32_t t = 0;
        for (int v = 0; v < SC_BS / 64; ++v) t += wcnt[v][threadIdx.x];
        block_count[int64_t(blockIdx.x) * SC_NIRREP + threadIdx.x] = t;
    }
}

__global__ void k_string_rank_offset(int64_t nstr, const uint8_t* __restrict__ irrep,
                                     const uint32_t* __restrict__ block_off, uint32_t* __restrict__ rank) {
    const int64_t I = int64_t(blockIdx.x) * SC_BS + threadIdx.x;
    if (I < nstr) rank[I] += block_off[int64_t(blockIdx.x) * SC_NIRREP + irrep[I]];
}

void string_classes_device(const SpinTables& t, const std::vector<int>& orbsym, uint8_t* d_irrep, uint32_t* d_rank,
                           std::vector<int64_t>& count, hipStream_t st) {
    if (int(orbsym.size()) != t.n) throw Error(-1, "orbital symmetry list has the wrong length");
    uint32_t bits[3] = {0, 0, 0};
    for (int p = 0; p < t.n; ++p) {
        if (orbsym[p] < 0 || orbsym[p] >= SC_NIRREP) throw Error(-1, "orbital irreps must be in [0, 8)");
        for (int b = 0; b < 3; ++b)
            if ((orbsym[p] >> b) & 1) bits[b] |= 1u << p;
    }
    count.assign(SC_NIRREP, 0);
    if (t.nstr == 0) return;
    const unsigned nb = unsigned((t.nstr + SC_BS - 1) / SC_BS);
    DBuf<uint32_t> dbits(3), bc(size_t(nb) * SC_NIRREP);
    dbits.upload(bits, 3);
    k_string_irrep<<<nb, SC_BS, 0, st>>>(t.str.p, t.nstr, dbits.p, d_irrep, d_rank, bc.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    std::vector<uint32_t> h = bc.to_host(), off(h.size());
    for (int g = 0; g < SC_NIRREP; ++g) {
        uint32_t run = 0;
        for (unsigned b = 0; b < nb; ++b) {
            off[size_t(b) * SC_NIRREP + g] = run;
            run += h[size_t(b) * SC_NIRREP + g];
        }
        count[g] = run;
    }
    DBuf<uint32_t> doff(off.size());
    doff.upload(off.data(), off.size());
    k_string_rank_offset<<<nb, SC_BS, 0, st>>>(t.nstr, d_irrep, doff.p, d_rank);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
}

// ---- Forte-order lists ---------------------------------------------------------------------
// Insert the reduced-space bits b (n' bits) into the orbitals not in `skip` (ascending).
__device__ __forceinline__ uint32_t d_expand(uint32_t b, int n, uint32_t skip) {
    uint32_t s = 0;
    int kk = 0;
    for (int i = 0; i < n; ++i) {
        if ((skip >> i) & 1u) continue;
        if ((b >> kk) & 1u) s |= 1u << i;
        ++kk;
    }
    return s;
}

__global__ void k_forte_vo(int64_t cnt, int n, int nred, int kred, int p, int q, const uint32_t* B,
                           double* sign, uint32_t* I, uint32_t* J) {
    int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    uint32_t b = d_unrank(uint32_t(t), nred, kred, B);
    uint32_t base = d_expand(b, n, (1u << p) | (1u << q));
    uint32_t Is = (base & ~(1u << p)) | (1u << q);
    uint32_t Js = (base & ~(1u << q)) | (1u << p);
    sign[t] = d_between_parity(base, p, q) ? -1.0 : 1.0;
    I[t] = d_rank(Is, n, B);
    J[t] = d_rank(Js, n, B);
}

__global__ void k_forte_oo(int64_t cnt, int n, int p, int q, int kred, const uint32_t* B,
                           double* sign, uint32_t* I, uint32_t* J) {
    int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    uint32_t b = d_unrank(uint32_t(t), n - 2, kred, B);
    uint32_t s = d_expand(b, n, (1u << p) | (1u << q)) | (1u << p) | (1u << q);
    uint32_t a = d_rank(s, n, B);
    sign[t] = 1.0;
    I[t] = a;
    J[t] = a;
}

__global__ void k_forte_vvoo(int64_t cnt, int n, int nred, int kred, int p, int q, int r, int s,
                             const uint32_t* B, double* sign, uint32_t* I, uint32_t* J) {
    int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    uint32_t b = d_unrank(uint32_t(t), nred, kred, B);
    uint32_t skip = (1u << p) | (1u << q) | (1u << r) | (1u << s);
    uint32_t Is = d_expand(b, n, skip) | (1u << r) | (1u << s);
    // a+_p a+_q a_s a_r applied to I (fci_string_vvoo_list.cc:166-182)
    uint32_t Js = Is;
    uint32_t neg = __popc(Js & (((1u << r) - 1u) & ~((1u << s) - 1u))) & 1u;  // bits in [s, r)
    Js &= ~((1u << r) | (1u << s));
    Js |= 1u << q;
    neg ^= __popc(Js & ((1u << q) - 1u)) & 1u;
    Js |= 1u << p;
    neg ^= __popc(Js & ((1u << p) - 1u)) & 1u;
    sign[t] = neg ? -1.0 : 1.0;
    I[t] = d_rank(Is, n, B);
    J[t] = d_rank(Js, n, B);
}

static int64_t run_list(int64_t cnt, std::vector<double>* sign, std::vector<uint32_t>* I,
                        std::vector<uint32_t>* J, const std::function<void(double*, uint32_t*, uint32_t*)>& launch) {
    if (cnt <= 0) {
        if (sign) sign->clear();
        if (I) I->clear();
        if (J) J->clear();
        return 0;
    }
    if (!sign) return cnt;
    DBuf<double> ds(cnt);
    DBuf<uint32_t> dI(cnt), dJ(cnt);
    launch(ds.p, dI.p, dJ.p);
    HIPCHK(hipGetLastError());
    *sign = ds.to_host();
    *I = dI.to_host();
    *J = dJ.to_host();
    return cnt;
}

int64_t forte_vo_list(const SpinTables& t, const uint32_t* B, int p, int q, std::vector<double>* sign,
                      std::vector<uint32_t>* I, std::vector<uint32_t>* J) {
    int nred = t.n - 1 - (p == q ? 0 : 1), kred = t.k - 1;
    int64_t cnt = (kred >= 0 && kred <= nred) ? int64_t(binom_host(nred, kred)) : 0;
    return run_list(cnt, sign, I, J, [&](double* s, uint32_t* a, uint32_t* b) {
        k_forte_vo<<<nblk(cnt, 256), 256>>>(cnt, t.n, nred, kred, p, q, B, s, a, b);
    });
}

int64_t forte_oo_list(const SpinTables& t, const uint32_t* B, int p, int q, std::vector<double>* sign,
                      std::vector<uint32_t>* I, std::vector<uint32_t>* J) {
    int kred = t.k - 2;
    int64_t cnt = kred >= 0 ? int64_t(binom_host(t.n - 2, kred)) : 0;
    return run_list(cnt, sign, I, J, [&](double* s, uint32_t* a, uint32_t* b) {
        k_forte_oo<<<nblk(cnt, 256), 256>>>(cnt, t.n, p, q, kred, B, s, a, b);
    });
}

int64_t forte_vvoo_list(const SpinTables& t, const uint32_t* B, int p, int q, int r, int s,
                        std::vector<double>* sign, std::vector<uint32_t>* I, std::vector<uint32_t>* J) {
    int overlap = (p == r) || (p == s) || (q == r) || (q == s);
    int nred = t.n - 4 + overlap, kred = t.k - 2;
    int64_t cnt = (kred >= 0 && nred >= 0 && nred >= kred) ? int64_t(binom_host(nred, kred)) : 0;
    return run_list(cnt, sign, I, J, [&](double* sg, uint32_t* a, uint32_t* b) {
        k_forte_vvoo<<<nblk(cnt, 256), 256>>>(cnt, t.n, nred, kred, p, q, r, s, B, sg, a, b);
    });
}

// ---- H diagonal ------------------------------------------------------------------------------
__global__ void k_hdiag(const uint32_t* __restrict__ sa, const uint32_t* __restrict__ sb, int64_t nb,
                        int64_t ndet, int n, const double* __restrict__ oa, const double* __restrict__ ob,
                        const double* __restrict__ taa, const double* __restrict__ tab,
                        const double* __restrict__ tbb, double e_core, double* __restrict__ out) {
    __shared__ double dg_a[kMaxOrb], dg_b[kMaxOrb];
    __shared__ double jaa[kMaxOrb * kMaxOrb], jab[kMaxOrb * kMaxOrb], jbb[kMaxOrb * kMaxOrb];
    const int n2 = n * n, n3 = n2 * n;
    for (int i = threadIdx.x; i < n2; i += blockDim.x) {
        int p = i / n, q = i % n;
        size_t ix = size_t(p) * n3 + size_t(q) * n2 + size_t(p) * n + q;
        jaa[i] = taa[ix];
        jab[i] = tab[ix];
        jbb[i] = tbb[ix];
        if (q == 0) {
            dg_a[p] = oa[p * n + p];
            dg_b[p] = ob[p * n + p];
        }
    }
    __syncthreads();
    int64_t I = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (I >= ndet) return;
    uint32_t A = sa[I / nb], Bs = sb[I % nb];
    // ActiveSpaceIntegrals::energy (active_space_integrals.cc:206-245), same summation order
    double e = 0.0;
    uint32_t a = A;
    while (a) {
        int p = __ffs(a) - 1;
        a &= a - 1u;
        e += dg_a[p];
        uint32_t ac = a;
        while (ac) {
            int q = __ffs(ac) - 1;
            ac &= ac - 1u;
            e += jaa[p * n + q];
        }
        uint32_t bc = Bs;
        while (bc) {
            int q = __ffs(bc) - 1;
            bc &= bc - 1u;
            e += jab[p * n + q];
        }
    }
    uint32_t b = Bs;
    while (b) {
        int p = __ffs(b) - 1;
        b &= b - 1u;
        e += dg_b[p];
        uint32_t bc = b;
        while (bc) {
            int q = __ffs(bc) - 1;
            bc &= bc - 1u;
            e += jbb[p * n + q];
        }
    }
    out[I] = e_core + e;
}

void hdiag_device(const SpinTables& a, const SpinTables& b, const double* oa, const double* ob,
                  const double* taa, const double* tab, const double* tbb, double e_core,
                  double* out, hipStream_t st) {
    int64_t nd = a.nstr * b.nstr;
    k_hdiag<<<nblk(nd, 256), 256, 0, st>>>(a.str.p, b.str.p, b.nstr, nd, a.n, oa, ob, taa, tab, tbb,
                                          e_core, out);
    HIPCHK(hipGetLastError());
}

// ---- k smallest by (value, index) ------------------------------------------------------------
constexpr int kSelChunk = 2048;

__device__ __forceinline__ bool kv_less(double va, int64_t ia, double vb, int64_t ib) {
    return va < vb || (va == vb && ia < ib);
}

// Each block sorts a chunk of (value,index) pairs in LDS (bitonic) and keeps the first k.
__global__ void __launch_bounds__(1024) k_select(const double* __restrict__ v, const int64_t* __restrict__ idx_in,
                                                 int64_t n, int k, double* ov, int64_t* oi) {
    __shared__ double sv[kSelChunk];
    __shared__ int64_t si[kSelChunk];
    int64_t base = int64_t(blockIdx.x) * kSelChunk;
    for (int i = threadIdx.x; i < kSelChunk; i += blockDim.x) {
        int64_t g = base + i;
        if (g < n) {
            sv[i] = v[g];
            si[i] = idx_in ? idx_in[g] : g;
        } else {
            sv[i] = __longlong_as_double(0x7FF0000000000000LL);  // +inf
            si[i] = INT64_MAX;
        }
    }
    __syncthreads();
    for (int size = 2; size <= kSelChunk; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < kSelChunk; i += blockDim.x) {
                int j = i ^ stride;
                if (j > i) {
                    bool up = (i & size) == 0;
                    bool sw = up ? kv_less(sv[j], si[j], sv[i], si[i]) : kv_less(sv[i], si[i], sv[j], si[j]);
                    if (sw) {
                        double tv = sv[i];
                        sv[i] = sv[j];
                        sv[j] = tv;
                        int64_t ti = si[i];
                        si[i] = si[j];
                        si[j] = ti;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < k; i += blockDim.x) {
        ov[int64_t(blockIdx.x) * k + i] = sv[i];
        oi[int64_t(blockIdx.x) * k + i] = si[i];
    }
}

void smallest_k(const double* d_v, int64_t n, int k, std::vector<double>& vals,
                std::vector<int64_t>& idx, hipStream_t st, SelectScratch* ws) {
    if (k > kSelChunk / 2) throw Error(-1, "smallest_k: k too large");
    if (k > n) k = int(n);
    SelectScratch local;
    SelectScratch& w = ws ? *ws : local;
    const double* cur_v = d_v;
    const int64_t* cur_i = nullptr;
    int64_t cur_n = n;
    int which = 0;
    while (true) {
        const int64_t nb = (cur_n + kSelChunk - 1) / kSelChunk;
        if (w.v[which].n < size_t(nb * k) || w.i[which].n < size_t(nb * k)) {
            HIPCHK(hipStreamSynchronize(st));  // the buffer about to be reallocated may still be read
            w.v[which].alloc(nb * k);
            w.i[which].alloc(nb * k);
        }
        k_select<<<unsigned(nb), 1024, 0, st>>>(cur_v, cur_i, cur_n, k, w.v[which].p, w.i[which].p);
        HIPCHK(hipGetLastError());
        cur_v = w.v[which].p;
        cur_i = w.i[which].p;
        cur_n = nb * k;
        if (nb == 1) break;
        which ^= 1;
    }
    HIPCHK(hipStreamSynchronize(st));
    vals.resize(k);
    idx.resize(k);
    HIPCHK(hipMemcpy(vals.data(), cur_v, k * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(idx.data(), cur_i, k * sizeof(int64_t), hipMemcpyDeviceToHost));
}

}  // namespace fci
