/*
 * fci_oracle.c -- CPU restatement of Forte's FCI hot path. Test infrastructure only: the
 * product never links this file. Citations are to forte 0.4.3 (paths under forte/).
 */
#include "fci_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static char g_err[512];
const char* orc_last_error(void) { return g_err; }
static void set_err(const char* msg) { snprintf(g_err, sizeof(g_err), "%s", msg); }

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static int popc(uint64_t x) { return __builtin_popcountll(x); }
/* (-1)^(number of set bits strictly below n): BitArray::slater_sign(n), bitarray.hpp:635 */
static double sign_below(uint64_t x, int n) {
    uint64_t mask = (n >= 64) ? ~0ULL : ((1ULL << n) - 1ULL);
    return (popc(x & mask) & 1) ? -1.0 : 1.0;
}
/* (-1)^(set bits strictly between m and n): ui64_sign(x, m, n), bitwise_operations.hpp:125 */
static double sign_between(uint64_t x, int m, int n) {
    if (m > n) {
        int t = m;
        m = n;
        n = t;
    }
    if (n - m < 2) return 1.0;
    uint64_t mask = ((1ULL << n) - 1ULL) & ~((1ULL << (m + 1)) - 1ULL);
    return (popc(x & mask) & 1) ? -1.0 : 1.0;
}

/* std::next_permutation on a sequence of 0/1 values (lexicographic, false < true). */
static int next_perm(int8_t* a, int n) {
    if (n < 2) return 0;
    int i = n - 2;
    while (i >= 0 && !(a[i] < a[i + 1])) --i;
    if (i < 0) {
        for (int l = 0, r = n - 1; l < r; ++l, --r) {
            int8_t t = a[l];
            a[l] = a[r];
            a[r] = t;
        }
        return 0;
    }
    int j = n - 1;
    while (!(a[i] < a[j])) --j;
    int8_t t = a[i];
    a[i] = a[j];
    a[j] = t;
    for (int l = i + 1, r = n - 1; l < r; ++l, --r) {
        t = a[l];
        a[l] = a[r];
        a[r] = t;
    }
    return 1;
}

uint64_t orc_binom(int n, int k) {
    if (k < 0 || k > n) return 0;
    uint64_t r = 1;
    for (int i = 1; i <= k; ++i) r = r * (uint64_t)(n - k + i) / (uint64_t)i;
    return r;
}

/* FCIStringLists::make_fci_strings, fci_string_lists.cc:258-293 (one irrep). */
int64_t orc_make_strings(int n, int k, uint64_t* out) {
    if (k < 0 || k > n || n > 63) return 0;
    int8_t b[64];
    for (int i = 0; i < n; ++i) b[i] = (i >= n - k) ? 1 : 0;
    int64_t cnt = 0;
    do {
        uint64_t s = 0;
        for (int i = 0; i < n; ++i)
            if (b[i]) s |= 1ULL << i;
        if (out) out[cnt] = s;
        ++cnt;
    } while (next_perm(b, n));
    return cnt;
}

/* ---------------------------------------------------------------------------------------- */
typedef struct {
    double* sign;
    uint32_t* I;
    uint32_t* J;
    int64_t size, cap;
} sublist;

static void sl_push(sublist* l, double s, uint32_t I, uint32_t J) {
    if (l->size == l->cap) {
        int64_t c = l->cap ? 2 * l->cap : 16;
        l->sign = (double*)realloc(l->sign, (size_t)c * sizeof(double));
        l->I = (uint32_t*)realloc(l->I, (size_t)c * sizeof(uint32_t));
        l->J = (uint32_t*)realloc(l->J, (size_t)c * sizeof(uint32_t));
        l->cap = c;
    }
    l->sign[l->size] = s;
    l->I[l->size] = I;
    l->J[l->size] = J;
    l->size++;
}
static void sl_free(sublist* l) {
    free(l->sign);
    free(l->I);
    free(l->J);
    memset(l, 0, sizeof(*l));
}

typedef struct {
    int k;
    int64_t nstr;
    uint64_t* str;
    int32_t* addr; /* table indexed by bit pattern, size 2^n */
    sublist* vo;   /* n*n, index p*n+q */
    sublist* oo;   /* npair */
    sublist* vvoo; /* npair*npair, index pq*npair+rs (pair indices) */
} spin_lists;

struct orc_ctx {
    int n, na, nb;
    int64_t npair;
    int* pair_p;
    int* pair_q;
    int32_t* pair_index; /* n*n -> pair index for p>q, else -1 */
    spin_lists sp[2];
    double *oei_a, *oei_b, *tei_aa, *tei_ab, *tei_bb;
    double e_core;
    double* CR; /* Forte's static temporaries, max_str x max_str */
    double* CL;
    int64_t maxstr;
    int lists; /* 0: strings and addresses only (orc_create_strings) */
    double timers[6];
};

static int64_t addr_of(const spin_lists* s, uint64_t str) { return s->addr[str]; }

/* FCIStringLists::make_vo, fci_string_vo_list.cc:68-140 */
static void make_vo(orc_ctx* ctx, spin_lists* s, int p, int q) {
    int n = ctx->n - 1 - (p == q ? 0 : 1);
    int k = s->k - 1;
    sublist* l = &s->vo[p * ctx->n + q];
    if (!(k >= 0 && k <= n)) return;
    int8_t b[64];
    for (int i = 0; i < n - k; ++i) b[i] = 0;
    for (int i = (n - k > 0 ? n - k : 0); i < n; ++i) b[i] = 1;
    int lo = p < q ? p : q, hi = p < q ? q : p;
    do {
        uint64_t I = 0;
        int kk = 0;
        double sign = 1.0;
        for (int i = 0; i < lo; ++i, ++kk)
            if (b[kk]) I |= 1ULL << i;
        for (int i = lo + 1; i < hi; ++i, ++kk)
            if (b[kk]) {
                I |= 1ULL << i;
                sign = -sign;
            }
        for (int i = hi + 1; i < ctx->n; ++i, ++kk)
            if (b[kk]) I |= 1ULL << i;
        uint64_t J = I;
        I &= ~(1ULL << p);
        I |= 1ULL << q;
        J &= ~(1ULL << q);
        J |= 1ULL << p;
        sl_push(l, sign, (uint32_t)addr_of(s, I), (uint32_t)addr_of(s, J));
    } while (next_perm(b, n));
}

/* FCIStringLists::make_oo, fci_string_oo_list.cc:59-115 */
static void make_oo(orc_ctx* ctx, spin_lists* s, int64_t pq) {
    int k = s->k - 2;
    if (k < 0) return;
    int p = ctx->pair_p[pq], q = ctx->pair_q[pq];
    int n = ctx->n - 2;
    int8_t b[64];
    for (int i = 0; i < n - k; ++i) b[i] = 0;
    for (int i = n - k; i < n; ++i) b[i] = 1;
    sublist* l = &s->oo[pq];
    do {
        uint64_t I = 0;
        int kk = 0;
        for (int i = 0; i < q; ++i, ++kk)
            if (b[kk]) I |= 1ULL << i;
        for (int i = q + 1; i < p; ++i, ++kk)
            if (b[kk]) I |= 1ULL << i;
        for (int i = p + 1; i < ctx->n; ++i, ++kk)
            if (b[kk]) I |= 1ULL << i;
        I |= (1ULL << p) | (1ULL << q);
        uint32_t a = (uint32_t)addr_of(s, I);
        sl_push(l, 1.0, a, a);
    } while (next_perm(b, n));
}

#define SWAP_IF(x, y)                                                                            \
    do {                                                                                         \
        if ((x) > (y)) {                                                                         \
            int t_ = (x);                                                                        \
            (x) = (y);                                                                           \
            (y) = t_;                                                                            \
        }                                                                                        \
    } while (0)

/* FCIStringLists::make_vvoo, fci_string_vvoo_list.cc:88-187 */
static void make_vvoo(orc_ctx* ctx, spin_lists* s, int p, int q, int r, int sidx) {
    int a[4] = {sidx, r, q, p};
    SWAP_IF(a[1], a[2]);
    SWAP_IF(a[0], a[1]);
    SWAP_IF(a[1], a[2]);
    SWAP_IF(a[2], a[3]);
    SWAP_IF(a[0], a[1]);
    SWAP_IF(a[1], a[2]);
    SWAP_IF(a[2], a[3]);
    int overlap = (a[0] == a[1]) || (a[1] == a[2]) || (a[2] == a[3]);
    int n = ctx->n - 4 + (overlap ? 1 : 0);
    int k = s->k - 2;
    if (!(k >= 0 && n >= 0 && n >= k)) return;
    int64_t pq = ctx->pair_index[p * ctx->n + q], rs = ctx->pair_index[r * ctx->n + sidx];
    sublist* l = &s->vvoo[pq * ctx->npair + rs];
    int8_t b[64];
    for (int i = 0; i < n - k; ++i) b[i] = 0;
    for (int i = n - k; i < n; ++i) b[i] = 1;
    do {
        uint64_t I = 0;
        I |= (1ULL << sidx) | (1ULL << r);
        int kk = 0;
        for (int i = 0; i < a[0]; ++i, ++kk)
            if (b[kk]) I |= 1ULL << i;
        for (int i = a[0] + 1; i < a[1]; ++i, ++kk)
            if (b[kk]) I |= 1ULL << i;
        for (int i = a[1] + 1; i < a[2]; ++i, ++kk)
            if (b[kk]) I |= 1ULL << i;
        for (int i = a[2] + 1; i < a[3]; ++i, ++kk)
            if (b[kk]) I |= 1ULL << i;
        for (int i = a[3] + 1; i < ctx->n; ++i, ++kk)
            if (b[kk]) I |= 1ULL << i;
        /* positions p,q not in {r,s} are empty in I by construction */
        uint64_t J = I;
        double sign = 1.0;
        for (int i = sidx; i < r; ++i)
            if ((J >> i) & 1ULL) sign = -sign;
        J &= ~((1ULL << r) | (1ULL << sidx));
        if (!((J >> q) & 1ULL)) {
            J |= 1ULL << q;
            sign *= sign_below(J, q);
            if (!((J >> p) & 1ULL)) {
                J |= 1ULL << p;
                sign *= sign_below(J, p);
                sl_push(l, sign, (uint32_t)addr_of(s, I), (uint32_t)addr_of(s, J));
            }
        }
    } while (next_perm(b, n));
}

static int build_spin(orc_ctx* ctx, spin_lists* s, int k, int with_lists) {
    int n = ctx->n;
    s->k = k;
    s->nstr = (int64_t)orc_binom(n, k);
    s->str = (uint64_t*)malloc((size_t)(s->nstr > 0 ? s->nstr : 1) * sizeof(uint64_t));
    orc_make_strings(n, k, s->str);
    s->addr = (int32_t*)malloc(sizeof(int32_t) << n);
    for (int64_t i = 0; i < (1LL << n); ++i) s->addr[i] = -1;
    for (int64_t i = 0; i < s->nstr; ++i) s->addr[s->str[i]] = (int32_t)i; /* fci_string_address.cc:45 */
    if (!with_lists) return 0;
    s->vo = (sublist*)calloc((size_t)(n * n), sizeof(sublist));
    s->oo = (sublist*)calloc((size_t)(ctx->npair > 0 ? ctx->npair : 1), sizeof(sublist));
    s->vvoo = (sublist*)calloc((size_t)(ctx->npair > 0 ? ctx->npair * ctx->npair : 1),
                               sizeof(sublist));
    /* make_vo_list: loop p, q (fci_string_vo_list.cc:68-84) */
    for (int p = 0; p < n; ++p)
        for (int q = 0; q < n; ++q) make_vo(ctx, s, p, q);
    for (int64_t pq = 0; pq < ctx->npair; ++pq) make_oo(ctx, s, pq);
    /* make_vvoo_list (fci_string_vvoo_list.cc:56-86) */
    for (int p = 0; p < n; ++p)
        for (int q = 0; q < n; ++q)
            for (int r = 0; r < n; ++r)
                for (int ss = 0; ss < n; ++ss)
                    if (p > q && r > ss && !(p == r && q == ss)) make_vvoo(ctx, s, p, q, r, ss);
    return 0;
}

static orc_ctx* create_ctx(int n, int na, int nb, const double* oei_a, const double* oei_b,
                           const double* tei_aa, const double* tei_ab, const double* tei_bb,
                           double e_core, int with_lists) {
    if (n < 1 || n > 24 || na < 0 || nb < 0 || na > n || nb > n) {
        set_err("orc_create: unsupported sizes (1 <= n <= 24, 0 <= na,nb <= n)");
        return NULL;
    }
    orc_ctx* ctx = (orc_ctx*)calloc(1, sizeof(orc_ctx));
    ctx->n = n;
    ctx->na = na;
    ctx->nb = nb;
    /* make_pair_list, fci_string_lists.cc:211-229 */
    ctx->npair = (int64_t)n * (n - 1) / 2;
    ctx->pair_p = (int*)malloc(sizeof(int) * (size_t)(ctx->npair + 1));
    ctx->pair_q = (int*)malloc(sizeof(int) * (size_t)(ctx->npair + 1));
    ctx->pair_index = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n * n));
    for (int i = 0; i < n * n; ++i) ctx->pair_index[i] = -1;
    int64_t c = 0;
    for (int p = 0; p < n; ++p)
        for (int q = 0; q < n; ++q)
            if (p > q) {
                ctx->pair_p[c] = p;
                ctx->pair_q[c] = q;
                ctx->pair_index[p * n + q] = (int32_t)c;
                ++c;
            }
    size_t n2 = (size_t)n * n, n4 = n2 * n2;
    ctx->oei_a = (double*)malloc(n2 * sizeof(double));
    ctx->oei_b = (double*)malloc(n2 * sizeof(double));
    ctx->tei_aa = (double*)malloc(n4 * sizeof(double));
    ctx->tei_ab = (double*)malloc(n4 * sizeof(double));
    ctx->tei_bb = (double*)malloc(n4 * sizeof(double));
    memcpy(ctx->oei_a, oei_a, n2 * sizeof(double));
    memcpy(ctx->oei_b, oei_b, n2 * sizeof(double));
    memcpy(ctx->tei_aa, tei_aa, n4 * sizeof(double));
    memcpy(ctx->tei_ab, tei_ab, n4 * sizeof(double));
    memcpy(ctx->tei_bb, tei_bb, n4 * sizeof(double));
    ctx->e_core = e_core;
    ctx->lists = with_lists;
    build_spin(ctx, &ctx->sp[0], na, with_lists);
    build_spin(ctx, &ctx->sp[1], nb, with_lists);
    ctx->maxstr = ctx->sp[0].nstr > ctx->sp[1].nstr ? ctx->sp[0].nstr : ctx->sp[1].nstr;
    if (with_lists) {
        /* FCIVector::allocate_temp_space, fci_vector.cc:54-79 */
        ctx->CR = (double*)calloc((size_t)(ctx->maxstr * ctx->maxstr), sizeof(double));
        ctx->CL = (double*)calloc((size_t)(ctx->maxstr * ctx->maxstr), sizeof(double));
    }
    return ctx;
}

orc_ctx* orc_create(int n, int na, int nb, const double* oei_a, const double* oei_b,
                    const double* tei_aa, const double* tei_ab, const double* tei_bb,
                    double e_core) {
    return create_ctx(n, na, nb, oei_a, oei_b, tei_aa, tei_ab, tei_bb, e_core, 1);
}

orc_ctx* orc_create_strings(int n, int na, int nb, const double* oei_a, const double* oei_b,
                            const double* tei_aa, const double* tei_ab, const double* tei_bb,
                            double e_core) {
    return create_ctx(n, na, nb, oei_a, oei_b, tei_aa, tei_ab, tei_bb, e_core, 0);
}

int orc_has_lists(const orc_ctx* ctx) { return ctx->lists; }

void orc_destroy(orc_ctx* ctx) {
    if (!ctx) return;
    for (int sp = 0; sp < 2; ++sp) {
        spin_lists* s = &ctx->sp[sp];
        if (s->vo)
            for (int i = 0; i < ctx->n * ctx->n; ++i) sl_free(&s->vo[i]);
        if (s->oo)
            for (int64_t i = 0; i < ctx->npair; ++i) sl_free(&s->oo[i]);
        if (s->vvoo)
            for (int64_t i = 0; i < ctx->npair * ctx->npair; ++i) sl_free(&s->vvoo[i]);
        free(s->vo);
        free(s->oo);
        free(s->vvoo);
        free(s->str);
        free(s->addr);
    }
    free(ctx->pair_p);
    free(ctx->pair_q);
    free(ctx->pair_index);
    free(ctx->oei_a);
    free(ctx->oei_b);
    free(ctx->tei_aa);
    free(ctx->tei_ab);
    free(ctx->tei_bb);
    free(ctx->CR);
    free(ctx->CL);
    free(ctx);
}

int64_t orc_nstr(const orc_ctx* ctx, int spin) { return ctx->sp[spin].nstr; }
int64_t orc_ndet(const orc_ctx* ctx) { return ctx->sp[0].nstr * ctx->sp[1].nstr; }
void orc_strings(const orc_ctx* ctx, int spin, uint64_t* out) {
    memcpy(out, ctx->sp[spin].str, (size_t)ctx->sp[spin].nstr * sizeof(uint64_t));
}
int64_t orc_address(const orc_ctx* ctx, int spin, uint64_t str) {
    if (str >> ctx->n) return -1;
    return ctx->sp[spin].addr[str];
}
static void sl_copy(const sublist* l, double* sign, uint32_t* I, uint32_t* J) {
    if (!l->size) return;
    memcpy(sign, l->sign, (size_t)l->size * sizeof(double));
    memcpy(I, l->I, (size_t)l->size * sizeof(uint32_t));
    memcpy(J, l->J, (size_t)l->size * sizeof(uint32_t));
}
int64_t orc_vo_size(const orc_ctx* ctx, int spin, int p, int q) {
    return ctx->sp[spin].vo[p * ctx->n + q].size;
}
void orc_vo_get(const orc_ctx* ctx, int spin, int p, int q, double* sign, uint32_t* I,
                uint32_t* J) {
    sl_copy(&ctx->sp[spin].vo[p * ctx->n + q], sign, I, J);
}
int64_t orc_npair(const orc_ctx* ctx) { return ctx->npair; }
void orc_pair(const orc_ctx* ctx, int64_t pq, int* p, int* q) {
    *p = ctx->pair_p[pq];
    *q = ctx->pair_q[pq];
}
int64_t orc_oo_size(const orc_ctx* ctx, int spin, int64_t pq) { return ctx->sp[spin].oo[pq].size; }
void orc_oo_get(const orc_ctx* ctx, int spin, int64_t pq, double* sign, uint32_t* I, uint32_t* J) {
    sl_copy(&ctx->sp[spin].oo[pq], sign, I, J);
}
static const sublist* vvoo_of(const orc_ctx* ctx, int spin, int p, int q, int r, int s) {
    int32_t pq = ctx->pair_index[p * ctx->n + q], rs = ctx->pair_index[r * ctx->n + s];
    if (pq < 0 || rs < 0) return NULL;
    return &ctx->sp[spin].vvoo[(int64_t)pq * ctx->npair + rs];
}
int64_t orc_vvoo_size(const orc_ctx* ctx, int spin, int p, int q, int r, int s) {
    const sublist* l = vvoo_of(ctx, spin, p, q, r, s);
    return l ? l->size : 0;
}
void orc_vvoo_get(const orc_ctx* ctx, int spin, int p, int q, int r, int s, double* sign,
                  uint32_t* I, uint32_t* J) {
    const sublist* l = vvoo_of(ctx, spin, p, q, r, s);
    if (l) sl_copy(l, sign, I, J);
}

/* ---------------------------------------------------------------------------------------- */
/* sigma (fci_vector_hamiltonian.cc). C is N_a x N_b row-major (fci_vector.cc:108-114).        */
static void daxpy(int64_t n, double a, const double* x, double* y) {
    for (int64_t i = 0; i < n; ++i) y[i] += a * x[i];
}
static double ddot(int64_t n, const double* x, const double* y) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += x[i] * y[i];
    return s;
}
#define IDX4(p, q, r, s, n) ((((size_t)(p) * (n) + (q)) * (n) + (r)) * (n) + (s))

/* gather_C_block / scatter_C_block, fci_vector.cc:239-278. For alpha the block is the vector
 * itself; for beta it is the transpose, held in the static CR / CL matrices. */
static void transpose_into(const double* c, double* m, int64_t na, int64_t nb) {
    for (int64_t a = 0; a < na; ++a)
        for (int64_t b = 0; b < nb; ++b) m[b * na + a] = c[a * nb + b];
}

static int g_stride = 1;  /* outer-loop sampling stride (1 = full sigma) */
static double g_done = 0.0, g_total = 0.0;

static void H1(orc_ctx* ctx, const double* C, double* R, int alfa) {
    int64_t na = ctx->sp[0].nstr, nb = ctx->sp[1].nstr;
    const double* Cr;
    double* Cl;
    if (alfa) {
        Cr = C;
        Cl = R;
    } else {
        transpose_into(C, ctx->CR, na, nb);
        memset(ctx->CL, 0, (size_t)(na * nb) * sizeof(double));
        Cr = ctx->CR;
        Cl = ctx->CL;
    }
    int64_t maxL = alfa ? nb : na;
    const spin_lists* s = &ctx->sp[alfa ? 0 : 1];
    const double* h = alfa ? ctx->oei_a : ctx->oei_b;
    int n = ctx->n;
    for (int p = 0; p < n; ++p)
        for (int q = 0; q < n; ++q) {
            const double Hpq = h[p * n + q];
            const sublist* l = &s->vo[p * n + q];
            g_total += (double)l->size;
            if ((p * n + q) % g_stride) continue;
            g_done += (double)l->size;
            for (int64_t e = 0; e < l->size; ++e)
                daxpy(maxL, l->sign[e] * Hpq, Cr + (int64_t)l->I[e] * maxL,
                      Cl + (int64_t)l->J[e] * maxL);
        }
    if (!alfa)
        for (int64_t a = 0; a < na; ++a)
            for (int64_t b = 0; b < nb; ++b) R[a * nb + b] += Cl[b * na + a];
}

static void H2_aaaa2(orc_ctx* ctx, const double* C, double* R, int alfa) {
    int64_t na = ctx->sp[0].nstr, nb = ctx->sp[1].nstr;
    const double* Cr;
    double* Cl;
    if (alfa) {
        Cr = C;
        Cl = R;
    } else {
        transpose_into(C, ctx->CR, na, nb);
        memset(ctx->CL, 0, (size_t)(na * nb) * sizeof(double));
        Cr = ctx->CR;
        Cl = ctx->CL;
    }
    int64_t maxL = alfa ? nb : na;
    const spin_lists* s = &ctx->sp[alfa ? 0 : 1];
    const double* tei = alfa ? ctx->tei_aa : ctx->tei_bb;
    int n = ctx->n;
    for (int64_t pq = 0; pq < ctx->npair; ++pq) {
        int p = ctx->pair_p[pq], q = ctx->pair_q[pq];
        const double integral = tei[IDX4(p, q, p, q, n)];
        const sublist* l = &s->oo[pq];
        g_total += (double)l->size;
        if (pq % g_stride) continue;
        g_done += (double)l->size;
        for (int64_t e = 0; e < l->size; ++e)
            daxpy(maxL, l->sign[e] * integral, Cr + (int64_t)l->I[e] * maxL,
                  Cl + (int64_t)l->J[e] * maxL);
    }
    for (int64_t pq = 0; pq < ctx->npair; ++pq) {
        int p = ctx->pair_p[pq], q = ctx->pair_q[pq];
        for (int64_t rs = 0; rs < pq; ++rs) {
            int r = ctx->pair_p[rs], ss = ctx->pair_q[rs];
            const double integral = tei[IDX4(p, q, r, ss, n)];
            const sublist* l1 = &s->vvoo[pq * ctx->npair + rs];
            const sublist* l2s = &s->vvoo[rs * ctx->npair + pq];
            g_total += (double)(l1->size + l2s->size);
            if (pq % g_stride) continue;
            g_done += (double)(l1->size + l2s->size);
            for (int64_t e = 0; e < l1->size; ++e)
                daxpy(maxL, l1->sign[e] * integral, Cr + (int64_t)l1->I[e] * maxL,
                      Cl + (int64_t)l1->J[e] * maxL);
            const sublist* l2 = &s->vvoo[rs * ctx->npair + pq];
            for (int64_t e = 0; e < l2->size; ++e)
                daxpy(maxL, l2->sign[e] * integral, Cr + (int64_t)l2->I[e] * maxL,
                      Cl + (int64_t)l2->J[e] * maxL);
        }
    }
    if (!alfa)
        for (int64_t a = 0; a < na; ++a)
            for (int64_t b = 0; b < nb; ++b) R[a * nb + b] += Cl[b * na + a];
}

static void H2_aabb(orc_ctx* ctx, const double* C, double* R, int rs_begin, int rs_end) {
    int64_t maxIa = ctx->sp[0].nstr, nb = ctx->sp[1].nstr;
    int n = ctx->n;
    int64_t ld = ctx->maxstr; /* CR / CL are max_str x max_str (fci_vector.cc:72-73) */
    for (int r = 0; r < n; ++r)
        for (int s = 0; s < n; ++s) {
            int rs_lin = r * n + s;
            if (rs_lin < rs_begin || rs_lin >= rs_end) continue;
            const sublist* vb = &ctx->sp[1].vo[r * n + s];
            int64_t maxSSb = vb->size;
            g_total += (double)maxSSb;
            if (rs_lin % g_stride) continue;
            g_done += (double)maxSSb;
            if (maxSSb == 0) continue;
            memset(ctx->CR, 0, (size_t)(ld * ld) * sizeof(double));
            memset(ctx->CL, 0, (size_t)(ld * ld) * sizeof(double));
            for (int64_t Ia = 0; Ia < maxIa; ++Ia) {
                const double* c = C + Ia * nb;
                double* cr = ctx->CR + Ia * ld;
                for (int64_t SSb = 0; SSb < maxSSb; ++SSb) cr[SSb] = c[vb->I[SSb]] * vb->sign[SSb];
            }
            for (int p = 0; p < n; ++p)
                for (int q = 0; q < n; ++q) {
                    const double integral = ctx->tei_ab[IDX4(p, r, q, s, n)];
                    const sublist* va = &ctx->sp[0].vo[p * n + q];
                    for (int64_t e = 0; e < va->size; ++e)
                        daxpy(maxSSb, integral * va->sign[e], ctx->CR + (int64_t)va->I[e] * ld,
                              ctx->CL + (int64_t)va->J[e] * ld);
                }
            for (int64_t Ja = 0; Ja < maxIa; ++Ja) {
                double* hc = R + Ja * nb;
                const double* cl = ctx->CL + Ja * ld;
                for (int64_t SSb = 0; SSb < maxSSb; ++SSb) hc[vb->J[SSb]] += cl[SSb];
            }
        }
}

void orc_sigma(orc_ctx* ctx, const double* C, double* sigma, int mask) {
    int64_t nd = orc_ndet(ctx);
    double t;
    memset(sigma, 0, (size_t)nd * sizeof(double));
    for (int i = 0; i < 6; ++i) ctx->timers[i] = 0.0;
    if (mask & ORC_H0) { /* H0, fci_vector_hamiltonian.cc:84-91 */
        t = now_s();
        for (int64_t i = 0; i < nd; ++i) sigma[i] = C[i] * ctx->e_core;
        ctx->timers[0] = now_s() - t;
    }
    if (mask & ORC_H1A) {
        t = now_s();
        H1(ctx, C, sigma, 1);
        ctx->timers[1] = now_s() - t;
    }
    if (mask & ORC_H1B) {
        t = now_s();
        H1(ctx, C, sigma, 0);
        ctx->timers[2] = now_s() - t;
    }
    if (mask & ORC_H2AABB) {
        t = now_s();
        H2_aabb(ctx, C, sigma, 0, ctx->n * ctx->n);
        ctx->timers[3] = now_s() - t;
    }
    if (mask & ORC_H2AAAA) {
        t = now_s();
        H2_aaaa2(ctx, C, sigma, 1);
        ctx->timers[4] = now_s() - t;
    }
    if (mask & ORC_H2BBBB) {
        t = now_s();
        H2_aaaa2(ctx, C, sigma, 0);
        ctx->timers[5] = now_s() - t;
    }
}

void orc_sigma_sample(orc_ctx* ctx, const double* C, double* sigma, int stride, double* t6, double* w6) {
    static const int masks[6] = {ORC_H0, ORC_H1A, ORC_H1B, ORC_H2AABB, ORC_H2AAAA, ORC_H2BBBB};
    int64_t nd = orc_ndet(ctx);
    memset(sigma, 0, (size_t)nd * sizeof(double));
    g_stride = stride < 1 ? 1 : stride;
    for (int t = 0; t < 6; ++t) {
        g_done = g_total = 0.0;
        double t0 = now_s();
        switch (masks[t]) {
            case ORC_H0:
                for (int64_t i = 0; i < nd; ++i) sigma[i] = C[i] * ctx->e_core;
                g_done = g_total = 1.0;
                break;
            case ORC_H1A: H1(ctx, C, sigma, 1); break;
            case ORC_H1B: H1(ctx, C, sigma, 0); break;
            case ORC_H2AABB: H2_aabb(ctx, C, sigma, 0, ctx->n * ctx->n); break;
            case ORC_H2AAAA: H2_aaaa2(ctx, C, sigma, 1); break;
            case ORC_H2BBBB: H2_aaaa2(ctx, C, sigma, 0); break;
        }
        t6[t] = now_s() - t0;
        w6[t] = g_total > 0.0 ? g_done / g_total : 1.0;
    }
    g_stride = 1;
}

void orc_sigma_timers(const orc_ctx* ctx, double* t6) { memcpy(t6, ctx->timers, 6 * sizeof(double)); }

void orc_sigma_aabb_range(orc_ctx* ctx, const double* C, double* sigma, int rs_begin, int rs_end) {
    H2_aabb(ctx, C, sigma, rs_begin, rs_end);
}

/* ActiveSpaceIntegrals::energy, active_space_integrals.cc:206-245 (e_core folded in; the
 * reference adds frozen_core here and nuclear+scalar in form_Hdiag_det). */
static double det_energy(const orc_ctx* ctx, uint64_t Ia, uint64_t Ib) {
    int n = ctx->n;
    size_t n3 = (size_t)n * n * n, n2 = (size_t)n * n;
    double energy = 0.0;
    uint64_t a = Ia;
    while (a) {
        int p = __builtin_ctzll(a);
        a &= a - 1;
        energy += ctx->oei_a[p * n + p];
        uint64_t ac = a;
        while (ac) {
            int q = __builtin_ctzll(ac);
            ac &= ac - 1;
            energy += ctx->tei_aa[p * n3 + q * n2 + p * n + q];
        }
        uint64_t bc = Ib;
        while (bc) {
            int q = __builtin_ctzll(bc);
            bc &= bc - 1;
            energy += ctx->tei_ab[p * n3 + q * n2 + p * n + q];
        }
    }
    uint64_t b = Ib;
    while (b) {
        int p = __builtin_ctzll(b);
        b &= b - 1;
        energy += ctx->oei_b[p * n + p];
        uint64_t bc = b;
        while (bc) {
            int q = __builtin_ctzll(bc);
            bc &= bc - 1;
            energy += ctx->tei_bb[p * n3 + q * n2 + p * n + q];
        }
    }
    return energy;
}

/* FCISolver::form_Hdiag_det, fci_solver_initial_guess.cc:151-172 */
void orc_hdiag(const orc_ctx* ctx, double* out) {
    int64_t na = ctx->sp[0].nstr, nb = ctx->sp[1].nstr;
    for (int64_t a = 0; a < na; ++a)
        for (int64_t b = 0; b < nb; ++b)
            out[a * nb + b] = ctx->e_core + det_energy(ctx, ctx->sp[0].str[a], ctx->sp[1].str[b]);
}

/* FCIVector::compute_spin2, fci_vector_spin2.cc:38-75 */
double orc_spin2(const orc_ctx* ctx, const double* C) {
    int n = ctx->n;
    int64_t nb = ctx->sp[1].nstr;
    double spin2 = 0.0;
#pragma omp parallel for collapse(2) schedule(dynamic) reduction(+ : spin2)
    for (int r = 0; r < n; ++r)
        for (int s = 0; s < n; ++s) {
            const sublist* va = &ctx->sp[0].vo[s * n + r];
            const sublist* vb = &ctx->sp[1].vo[r * n + s];
            for (int64_t ea = 0; ea < va->size; ++ea)
                for (int64_t eb = 0; eb < vb->size; ++eb)
                    spin2 += C[(int64_t)va->J[ea] * nb + vb->J[eb]] *
                             C[(int64_t)va->I[ea] * nb + vb->I[eb]] * va->sign[ea] * vb->sign[eb];
        }
    double na_ = ctx->na, nb_ = ctx->nb;
    return -spin2 + 0.25 * (na_ - nb_) * (na_ - nb_) + 0.5 * (na_ + nb_);
}

/* FCIVector::compute_rdms and helpers, fci_vector_rdm.cc:43-393 */
void orc_rdms(const orc_ctx* ctx, const double* CLv, const double* CRv, int level, double* g1a,
              double* g1b, double* g2aa, double* g2ab, double* g2bb) {
    int n = ctx->n;
    int64_t na = ctx->sp[0].nstr, nb = ctx->sp[1].nstr;
    size_t n2 = (size_t)n * n, n4 = n2 * n2;
    double* tL = (double*)malloc((size_t)(na * nb) * sizeof(double));
    double* tR = (double*)malloc((size_t)(na * nb) * sizeof(double));
    if (level >= 1) {
        for (int alfa = 1; alfa >= 0; --alfa) {
            double* rdm = alfa ? g1a : g1b;
            memset(rdm, 0, n2 * sizeof(double));
            if ((alfa && ctx->na < 1) || (!alfa && ctx->nb < 1)) continue;
            const double *Cl = CLv, *Cr = CRv;
            if (!alfa) {
                transpose_into(CLv, tL, na, nb);
                transpose_into(CRv, tR, na, nb);
                Cl = tL;
                Cr = tR;
            }
            int64_t maxL = alfa ? nb : na;
            const spin_lists* s = &ctx->sp[alfa ? 0 : 1];
#pragma omp parallel for collapse(2) schedule(dynamic)
            for (int p = 0; p < n; ++p)
                for (int q = 0; q < n; ++q) {
                    const sublist* l = &s->vo[p * n + q];
                    double el = 0.0;
                    for (int64_t e = 0; e < l->size; ++e)
                        el += l->sign[e] * ddot(maxL, Cl + (int64_t)l->J[e] * maxL,
                                                Cr + (int64_t)l->I[e] * maxL);
                    rdm[p * n + q] += el;
                }
        }
    }
    if (level >= 2) {
        for (int alfa = 1; alfa >= 0; --alfa) {
            double* rdm = alfa ? g2aa : g2bb;
            memset(rdm, 0, n4 * sizeof(double));
            if ((alfa && ctx->na < 2) || (!alfa && ctx->nb < 2)) continue;
            const double *Cl = CLv, *Cr = CRv;
            if (!alfa) {
                transpose_into(CLv, tL, na, nb);
                transpose_into(CRv, tR, na, nb);
                Cl = tL;
                Cr = tR;
            }
            int64_t maxL = alfa ? nb : na;
            const spin_lists* s = &ctx->sp[alfa ? 0 : 1];
            /* every (pq) iteration writes its own elements (the blocks pq and the transposed
             * ones), so the pair loops run in parallel without races */
#pragma omp parallel for schedule(dynamic)
            for (int64_t pq = 0; pq < ctx->npair; ++pq) {
                int p = ctx->pair_p[pq], q = ctx->pair_q[pq];
                const sublist* l = &s->oo[pq];
                double el = 0.0;
                for (int64_t e = 0; e < l->size; ++e)
                    el += l->sign[e] * ddot(maxL, Cl + (int64_t)l->J[e] * maxL,
                                            Cr + (int64_t)l->I[e] * maxL);
                rdm[IDX4(p, q, p, q, n)] += el;
                rdm[IDX4(p, q, q, p, n)] -= el;
                rdm[IDX4(q, p, p, q, n)] -= el;
                rdm[IDX4(q, p, q, p, n)] += el;
            }
#pragma omp parallel for schedule(dynamic)
            for (int64_t pq = 0; pq < ctx->npair; ++pq) {
                int p = ctx->pair_p[pq], q = ctx->pair_q[pq];
                for (int64_t rs = 0; rs < pq; ++rs) {
                    int r = ctx->pair_p[rs], ss = ctx->pair_q[rs];
                    const sublist* l = &s->vvoo[pq * ctx->npair + rs];
                    double el = 0.0;
                    for (int64_t e = 0; e < l->size; ++e)
                        el += l->sign[e] * ddot(maxL, Cl + (int64_t)l->J[e] * maxL,
                                                Cr + (int64_t)l->I[e] * maxL);
                    rdm[IDX4(p, q, r, ss, n)] += el;
                    rdm[IDX4(q, p, r, ss, n)] -= el;
                    rdm[IDX4(p, q, ss, r, n)] -= el;
                    rdm[IDX4(q, p, ss, r, n)] += el;
                    rdm[IDX4(r, ss, p, q, n)] += el;
                    rdm[IDX4(r, ss, q, p, n)] -= el;
                    rdm[IDX4(ss, r, p, q, n)] -= el;
                    rdm[IDX4(ss, r, q, p, n)] += el;
                }
            }
        }
        memset(g2ab, 0, n4 * sizeof(double));
        if (ctx->na >= 1 && ctx->nb >= 1) {
            /* each (r, s) writes the elements [p, r, q, s] of its own: parallel without races */
#pragma omp parallel for collapse(2) schedule(dynamic)
            for (int r = 0; r < n; ++r)
                for (int s = 0; s < n; ++s) {
                    const sublist* vb = &ctx->sp[1].vo[r * n + s];
                    for (int p = 0; p < n; ++p)
                        for (int q = 0; q < n; ++q) {
                            const sublist* va = &ctx->sp[0].vo[p * n + q];
                            double el = 0.0;
                            for (int64_t ea = 0; ea < va->size; ++ea)
                                for (int64_t eb = 0; eb < vb->size; ++eb)
                                    el += CLv[(int64_t)va->J[ea] * nb + vb->J[eb]] *
                                          CRv[(int64_t)va->I[ea] * nb + vb->I[eb]] * va->sign[ea] *
                                          vb->sign[eb];
                            g2ab[IDX4(p, r, q, s, n)] += el;
                        }
                }
        }
    }
    free(tL);
    free(tR);
}

/* ---------------------------------------------------------------------------------------- */
/* Slater rules, active_space_integrals.cc:247-411. Determinants as (alpha word, beta word). */
static double slater_sign_aaaa(uint64_t w, int i, int j, int a, int b) {
    /* DeterminantImpl::slater_sign_aaaa, determinant.hpp:390-409 */
    if ((((i < a) && (j < a) && (i < b) && (j < b)) == 1) ||
        (((i < a) || (j < a) || (i < b) || (j < b)) == 0)) {
        if ((i < j) ^ (a < b))
            return -1.0 * sign_between(w, i, j) * sign_between(w, a, b);
        return sign_between(w, i, j) * sign_between(w, a, b);
    }
    if ((i < j) ^ (a < b)) return -1.0 * sign_between(w, i, b) * sign_between(w, j, a);
    return sign_between(w, i, a) * sign_between(w, j, b);
}

static double slater_rules(const orc_ctx* ctx, uint64_t la, uint64_t lb, uint64_t ra, uint64_t rb) {
    int n = ctx->n;
    size_t n3 = (size_t)n * n * n, n2 = (size_t)n * n;
    if (popc(la) != popc(ra) || popc(lb) != popc(rb)) return 0.0;
    int nadiff = popc(la ^ ra), nbdiff = popc(lb ^ rb);
    if (nadiff + nbdiff > 4) return 0.0;
    nadiff /= 2;
    nbdiff /= 2;
    double me = 0.0;
    if (nadiff == 0 && nbdiff == 0) {
        for (int p = 0; p < n; ++p) {
            int ap = (la >> p) & 1, bp = (lb >> p) & 1;
            if (ap) me += ctx->oei_a[p * n + p];
            if (bp) me += ctx->oei_b[p * n + p];
            for (int q = 0; q < n; ++q) {
                int aq = (la >> q) & 1, bq = (lb >> q) & 1;
                if (ap && aq) me += 0.5 * ctx->tei_aa[p * n3 + q * n2 + p * n + q];
                if (bp && bq) me += 0.5 * ctx->tei_bb[p * n3 + q * n2 + p * n + q];
                if (ap && bq) me += ctx->tei_ab[p * n3 + q * n2 + p * n + q];
            }
        }
    }
    if (nadiff == 1 && nbdiff == 0) {
        int i = 0, j = 0;
        for (int p = 0; p < n; ++p) {
            if ((((la ^ ra) >> p) & 1) && ((la >> p) & 1)) i = p;
            if ((((la ^ ra) >> p) & 1) && ((ra >> p) & 1)) j = p;
        }
        double sign = sign_between(la, i, j);
        me = sign * ctx->oei_a[i * n + j];
        for (int p = 0; p < n; ++p) {
            if (((la >> p) & 1) && ((ra >> p) & 1)) me += sign * ctx->tei_aa[i * n3 + p * n2 + j * n + p];
            if (((lb >> p) & 1) && ((rb >> p) & 1)) me += sign * ctx->tei_ab[i * n3 + p * n2 + j * n + p];
        }
    }
    if (nadiff == 0 && nbdiff == 1) {
        int i = 0, j = 0;
        for (int p = 0; p < n; ++p) {
            if ((((lb ^ rb) >> p) & 1) && ((lb >> p) & 1)) i = p;
            if ((((lb ^ rb) >> p) & 1) && ((rb >> p) & 1)) j = p;
        }
        double sign = sign_between(lb, i, j);
        me = sign * ctx->oei_b[i * n + j];
        for (int p = 0; p < n; ++p) {
            if (((la >> p) & 1) && ((ra >> p) & 1)) me += sign * ctx->tei_ab[p * n3 + i * n2 + p * n + j];
            if (((lb >> p) & 1) && ((rb >> p) & 1)) me += sign * ctx->tei_bb[i * n3 + p * n2 + j * n + p];
        }
    }
    if (nadiff == 2 && nbdiff == 0) {
        int i = -1, j = 0, k = -1, l = 0;
        for (int p = 0; p < n; ++p) {
            if ((((la ^ ra) >> p) & 1) && ((la >> p) & 1)) {
                if (i == -1) i = p; else j = p;
            }
            if ((((la ^ ra) >> p) & 1) && ((ra >> p) & 1)) {
                if (k == -1) k = p; else l = p;
            }
        }
        me = slater_sign_aaaa(la, i, j, k, l) * ctx->tei_aa[i * n3 + j * n2 + k * n + l];
    }
    if (nadiff == 0 && nbdiff == 2) {
        int i = -1, j = -1, k = -1, l = -1;
        for (int p = 0; p < n; ++p) {
            if ((((lb ^ rb) >> p) & 1) && ((lb >> p) & 1)) {
                if (i == -1) i = p; else j = p;
            }
            if ((((lb ^ rb) >> p) & 1) && ((rb >> p) & 1)) {
                if (k == -1) k = p; else l = p;
            }
        }
        me = slater_sign_aaaa(lb, i, j, k, l) * ctx->tei_bb[i * n3 + j * n2 + k * n + l];
    }
    if (nadiff == 1 && nbdiff == 1) {
        int i = -1, j = -1, k = -1, l = -1;
        for (int p = 0; p < n; ++p) {
            if ((((la ^ ra) >> p) & 1) && ((la >> p) & 1)) i = p;
            if ((((lb ^ rb) >> p) & 1) && ((lb >> p) & 1)) j = p;
            if ((((la ^ ra) >> p) & 1) && ((ra >> p) & 1)) k = p;
            if ((((lb ^ rb) >> p) & 1) && ((rb >> p) & 1)) l = p;
        }
        double sign = sign_between(la, i, k) * sign_between(lb, j, l);
        me = sign * ctx->tei_ab[i * n3 + j * n2 + k * n + l];
    }
    return me;
}

void orc_dense_hamiltonian(const orc_ctx* ctx, double* H) {
    int64_t na = ctx->sp[0].nstr, nb = ctx->sp[1].nstr, nd = na * nb;
    for (int64_t I = 0; I < nd; ++I)
        for (int64_t J = I; J < nd; ++J) {
            double h = slater_rules(ctx, ctx->sp[0].str[I / nb], ctx->sp[1].str[I % nb],
                                    ctx->sp[0].str[J / nb], ctx->sp[1].str[J % nb]);
            if (I == J) h += ctx->e_core;
            H[I * nd + J] = h;
            H[J * nd + I] = h;
        }
}

/* The strings reached from w by at most two excitations (same electron count): level 0 is w
 * itself, level 1 the single, level 2 the double excitations. */
typedef struct {
    uint64_t* str;
    int* level;
    int64_t size;
} reach_t;

static reach_t reach_within_two(uint64_t w, int n) {
    uint64_t full = (n >= 64) ? ~0ull : ((1ull << n) - 1);
    uint64_t occ = w, vir = full & ~w;
    int no = popc(occ), nv = popc(vir);
    int64_t cap = 1 + (int64_t)no * nv + (int64_t)(no * (no - 1) / 2) * (nv * (nv - 1) / 2);
    reach_t r;
    r.str = (uint64_t*)malloc((size_t)cap * sizeof(uint64_t));
    r.level = (int*)malloc((size_t)cap * sizeof(int));
    r.size = 0;
    r.str[r.size] = w;
    r.level[r.size++] = 0;
    for (int i = 0; i < n; ++i) {
        if (!((occ >> i) & 1)) continue;
        for (int a = 0; a < n; ++a) {
            if (!((vir >> a) & 1)) continue;
            r.str[r.size] = w ^ (1ull << i) ^ (1ull << a);
            r.level[r.size++] = 1;
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            if (!((occ >> i) & 1) || !((occ >> j) & 1)) continue;
            for (int a = 0; a < n; ++a)
                for (int b = a + 1; b < n; ++b) {
                    if (!((vir >> a) & 1) || !((vir >> b) & 1)) continue;
                    r.str[r.size] = w ^ (1ull << i) ^ (1ull << j) ^ (1ull << a) ^ (1ull << b);
                    r.level[r.size++] = 2;
                }
        }
    return r;
}

/* Selected rows of sigma = H C by the Slater-Condon rules (slater_rules above,
 * active_space_integrals.cc:247-411), summed over every determinant connected to the row: the
 * diagonal, alpha and beta single excitations and alpha-alpha, beta-beta and alpha-beta double
 * excitations. Independent of the string lists, so it also checks sigma at sizes whose lists
 * or full sigma the oracle cannot afford. out[r] = (H C)[rows[r]]. */
void orc_sigma_rows(const orc_ctx* ctx, const double* C, int64_t nrows, const int64_t* rows,
                    double* out) {
    const int64_t nb = ctx->sp[1].nstr;
    const int n = ctx->n;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t r = 0; r < nrows; ++r) {
        const int64_t I = rows[r];
        const uint64_t la = ctx->sp[0].str[I / nb], lb = ctx->sp[1].str[I % nb];
        reach_t ra = reach_within_two(la, n), rb = reach_within_two(lb, n);
        double acc = 0.0;
        for (int64_t x = 0; x < ra.size; ++x) {
            const int64_t Ja = ctx->sp[0].addr[ra.str[x]];
            for (int64_t y = 0; y < rb.size; ++y) {
                if (ra.level[x] + rb.level[y] > 2) continue;
                const int64_t J = Ja * nb + ctx->sp[1].addr[rb.str[y]];
                double h = slater_rules(ctx, la, lb, ra.str[x], rb.str[y]);
                if (J == I) h += ctx->e_core;
                acc += h * C[J];
            }
        }
        out[r] = acc;
        free(ra.str);
        free(ra.level);
        free(rb.str);
        free(rb.level);
    }
}

/* ---- explicit second quantization on spin-orbital strings ------------------------------ */
/* Spin orbital index: alpha p -> p, beta p -> n + p (Determinant: alpha word first). */
typedef struct {
    uint64_t a, b;
} det2;
static int so_occ(det2 d, int n, int so) { return so < n ? (int)((d.a >> so) & 1) : (int)((d.b >> (so - n)) & 1); }
static int so_count_below(det2 d, int n, int so) {
    if (so < n) return popc(d.a & ((1ULL << so) - 1ULL));
    return popc(d.a) + popc(d.b & ((1ULL << (so - n)) - 1ULL));
}
static void so_flip(det2* d, int n, int so) {
    if (so < n) d->a ^= 1ULL << so; else d->b ^= 1ULL << (so - n);
}
/* apply ops right-to-left: ops[nop-1] acts first; create[i] = 1 for a+, 0 for a */
static double apply_ops(det2* d, int n, int nop, const int* so, const int* create) {
    double sign = 1.0;
    for (int i = nop - 1; i >= 0; --i) {
        int occ = so_occ(*d, n, so[i]);
        if (create[i] == occ) return 0.0;
        if (so_count_below(*d, n, so[i]) & 1) sign = -sign;
        so_flip(d, n, so[i]);
    }
    return sign;
}

static int64_t det_index(const orc_ctx* ctx, det2 d) {
    int64_t ia = orc_address(ctx, 0, d.a), ib = orc_address(ctx, 1, d.b);
    if (ia < 0 || ib < 0) return -1;
    return ia * ctx->sp[1].nstr + ib;
}

void orc_brute_sigma(const orc_ctx* ctx, const double* C, double* sigma) {
    int n = ctx->n;
    int64_t na = ctx->sp[0].nstr, nb = ctx->sp[1].nstr, nd = na * nb;
    for (int64_t I = 0; I < nd; ++I) sigma[I] = ctx->e_core * C[I];
    for (int64_t I = 0; I < nd; ++I) {
        if (C[I] == 0.0) continue;
        det2 d0 = {ctx->sp[0].str[I / nb], ctx->sp[1].str[I % nb]};
        /* one-body, both spins */
        for (int sp = 0; sp < 2; ++sp) {
            const double* h = sp ? ctx->oei_b : ctx->oei_a;
            for (int p = 0; p < n; ++p)
                for (int q = 0; q < n; ++q) {
                    det2 d = d0;
                    int so[2] = {p + sp * n, q + sp * n}, cr[2] = {1, 0};
                    double sg = apply_ops(&d, n, 2, so, cr);
                    if (sg == 0.0) continue;
                    int64_t J = det_index(ctx, d);
                    sigma[J] += sg * h[p * n + q] * C[I];
                }
        }
        /* two-body: 1/4 <pq||rs> a+p a+q a_s a_r (same spin), <pq|rs> a+pA a+qB a_sB a_rA */
        for (int p = 0; p < n; ++p)
            for (int q = 0; q < n; ++q)
                for (int r = 0; r < n; ++r)
                    for (int s = 0; s < n; ++s) {
                        for (int sp = 0; sp < 2; ++sp) {
                            const double* t = sp ? ctx->tei_bb : ctx->tei_aa;
                            double v = t[IDX4(p, q, r, s, n)];
                            if (v == 0.0) continue;
                            det2 d = d0;
                            int so[4] = {p + sp * n, q + sp * n, s + sp * n, r + sp * n};
                            int cr[4] = {1, 1, 0, 0};
                            double sg = apply_ops(&d, n, 4, so, cr);
                            if (sg == 0.0) continue;
                            sigma[det_index(ctx, d)] += 0.25 * sg * v * C[I];
                        }
                        double v = ctx->tei_ab[IDX4(p, q, r, s, n)];
                        if (v == 0.0) continue;
                        det2 d = d0;
                        int so[4] = {p, q + n, s + n, r};
                        int cr[4] = {1, 1, 0, 0};
                        double sg = apply_ops(&d, n, 4, so, cr);
                        if (sg == 0.0) continue;
                        sigma[det_index(ctx, d)] += sg * v * C[I];
                    }
    }
}

void orc_brute_rdms(const orc_ctx* ctx, const double* CL, const double* CR, double* g1a,
                    double* g1b, double* g2aa, double* g2ab, double* g2bb) {
    int n = ctx->n;
    int64_t na = ctx->sp[0].nstr, nb = ctx->sp[1].nstr, nd = na * nb;
    size_t n2 = (size_t)n * n, n4 = n2 * n2;
    memset(g1a, 0, n2 * sizeof(double));
    memset(g1b, 0, n2 * sizeof(double));
    memset(g2aa, 0, n4 * sizeof(double));
    memset(g2ab, 0, n4 * sizeof(double));
    memset(g2bb, 0, n4 * sizeof(double));
    for (int64_t I = 0; I < nd; ++I) {
        if (CR[I] == 0.0) continue;
        det2 d0 = {ctx->sp[0].str[I / nb], ctx->sp[1].str[I % nb]};
        for (int p = 0; p < n; ++p)
            for (int q = 0; q < n; ++q)
                for (int sp = 0; sp < 2; ++sp) {
                    det2 d = d0;
                    int so[2] = {p + sp * n, q + sp * n}, cr[2] = {1, 0};
                    double sg = apply_ops(&d, n, 2, so, cr);
                    if (sg == 0.0) continue;
                    (sp ? g1b : g1a)[p * n + q] += sg * CL[det_index(ctx, d)] * CR[I];
                }
        for (int p = 0; p < n; ++p)
            for (int q = 0; q < n; ++q)
                for (int r = 0; r < n; ++r)
                    for (int s = 0; s < n; ++s) {
                        /* <a+p a+q a_s a_r> */
                        for (int kind = 0; kind < 3; ++kind) {
                            int sp1 = kind == 2 ? n : 0, sp2 = kind == 0 ? 0 : n;
                            det2 d = d0;
                            int so[4] = {p + sp1, q + sp2, s + sp2, r + sp1};
                            int cr[4] = {1, 1, 0, 0};
                            double sg = apply_ops(&d, n, 4, so, cr);
                            if (sg == 0.0) continue;
                            double* g = kind == 0 ? g2aa : (kind == 1 ? g2ab : g2bb);
                            g[IDX4(p, q, r, s, n)] += sg * CL[det_index(ctx, d)] * CR[I];
                        }
                    }
    }
}

/* one operator (create or annihilate spin orbital so) on d; returns its sign, 0 if it vanishes */
static double apply1(det2* d, int n, int so, int create) {
    if (create == so_occ(*d, n, so)) return 0.0;
    double sign = (so_count_below(*d, n, so) & 1) ? -1.0 : 1.0;
    so_flip(d, n, so);
    return sign;
}

/* Brute-force 3-RDMs <CL| a+p a+q a+r a_u a_t a_s |CR> at [p,q,r,s,t,u] (n^6 each), the
 * definition the reference's own check uses (FCIVector::test_rdms, fci_vector_rdm.cc:808-845:
 * destroy s, t, u, then create r, q, p on the right determinant). Spin patterns: aaa; aab =
 * (p,q,s,t alpha; r,u beta); abb = (p,s alpha; q,r,t,u beta); bbb. Operators are applied one at a
 * time so that vanishing strings are pruned early. */
void orc_brute_rdm3(const orc_ctx* ctx, const double* CL, const double* CR, double* g3aaa, double* g3aab,
                    double* g3abb, double* g3bbb) {
    const int n = ctx->n;
    const int64_t nb = ctx->sp[1].nstr, nd = ctx->sp[0].nstr * nb;
    const size_t n6 = (size_t)n * n * n * n * n * n;
    double* g[4] = {g3aaa, g3aab, g3abb, g3bbb};
    /* spin of p, q, r, s, t, u (0 alpha, 1 beta) per pattern */
    static const int spin[4][6] = {{0, 0, 0, 0, 0, 0}, {0, 0, 1, 0, 0, 1}, {0, 1, 1, 0, 1, 1}, {1, 1, 1, 1, 1, 1}};
    for (int k = 0; k < 4; ++k) memset(g[k], 0, n6 * sizeof(double));
    for (int64_t I = 0; I < nd; ++I) {
        if (CR[I] == 0.0) continue;
        const det2 d0 = {ctx->sp[0].str[I / nb], ctx->sp[1].str[I % nb]};
        for (int k = 0; k < 4; ++k) {
            const int* sp = spin[k];
            for (int ss = 0; ss < n; ++ss) {
                det2 d1 = d0;
                double g1 = apply1(&d1, n, ss + sp[3] * n, 0);
                if (g1 == 0.0) continue;
                for (int t = 0; t < n; ++t) {
                    det2 d2 = d1;
                    double g2 = g1 * apply1(&d2, n, t + sp[4] * n, 0);
                    if (g2 == 0.0) continue;
                    for (int u = 0; u < n; ++u) {
                        det2 d3 = d2;
                        double g3 = g2 * apply1(&d3, n, u + sp[5] * n, 0);
                        if (g3 == 0.0) continue;
                        for (int r = 0; r < n; ++r) {
                            det2 d4 = d3;
                            double g4 = g3 * apply1(&d4, n, r + sp[2] * n, 1);
                            if (g4 == 0.0) continue;
                            for (int q = 0; q < n; ++q) {
                                det2 d5 = d4;
                                double g5 = g4 * apply1(&d5, n, q + sp[1] * n, 1);
                                if (g5 == 0.0) continue;
                                for (int pp = 0; pp < n; ++pp) {
                                    det2 d6 = d5;
                                    double g6 = g5 * apply1(&d6, n, pp + sp[0] * n, 1);
                                    if (g6 == 0.0) continue;
                                    const int64_t J = det_index(ctx, d6);
                                    if (J < 0) continue;
                                    const size_t idx = ((((((size_t)pp * n + q) * n + r) * n + ss) * n + t) * n + u);
                                    g[k][idx] += g6 * CL[J] * CR[I];
                                }
                            }
                        }
                    }
                }
            }
        }
    }
}

/* ---------------------------------------------------------------------------------------- */
/* Cyclic Jacobi for dense symmetric matrices; eigenvalues ascending. */
void orc_jacobi(int n, const double* A0, double* evals, double* evecs) {
    double* A = (double*)malloc((size_t)n * n * sizeof(double));
    double* V = (double*)malloc((size_t)n * n * sizeof(double));
    memcpy(A, A0, (size_t)n * n * sizeof(double));
    for (int i = 0; i < n * n; ++i) V[i] = 0.0;
    for (int i = 0; i < n; ++i) V[i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                tot += A[i * n + j] * A[i * n + j];
                if (i != j) off += A[i * n + j] * A[i * n + j];
            }
        if (off <= 1e-30 * (tot > 0 ? tot : 1.0) || off == 0.0) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = A[p * n + q];
                if (fabs(apq) < 1e-300) continue;
                double app = A[p * n + p], aqq = A[q * n + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = c * vkp - s * vkq;
                    V[k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    /* sort ascending (stable selection by index for ties) */
    int* idx = (int*)malloc((size_t)n * sizeof(int));
    for (int i = 0; i < n; ++i) idx[i] = i;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
            if (A[idx[j] * n + idx[j]] < A[idx[i] * n + idx[i]]) {
                int t = idx[i];
                idx[i] = idx[j];
                idx[j] = t;
            }
    for (int k = 0; k < n; ++k) {
        evals[k] = A[idx[k] * n + idx[k]];
        for (int i = 0; i < n; ++i) evecs[(size_t)k * n + i] = V[i * n + idx[k]];
    }
    free(idx);
    free(A);
    free(V);
}

/* ---------------------------------------------------------------------------------------- */
/* Davidson-Liu, helpers/davidson_liu_solver.cc. Vectors are rows of subspace x size arrays.  */
typedef void (*sigma_fn)(void* user, const double* b, double* sigma);

typedef struct {
    int64_t size;
    int nroot;
    int collapse_size, subspace_size;
    int basis_size, sigma_size;
    int maxiter;
    double e_conv, r_conv;
    double discard, orth;
    double *temp, *b, *r, *sigma;
    double *G, *alpha, *lambda, *lambda_old, *resid;
    int nproj;
    const int64_t *poff, *pidx;
    const double* pval;
    uint64_t rng;
} dl_t;

static uint64_t splitmix64(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static double rng_uniform01(uint64_t* s) { return ((double)(splitmix64(s) >> 11) + 0.5) * 0x1.0p-53; }

static int dl_init(dl_t* d, int64_t size, int nroot, int cpr, int spr) {
    memset(d, 0, sizeof(*d));
    if (size == 0) { set_err("DavidsonLiuSolver called with space of dimension zero."); return -1; }
    if (nroot == 0) { set_err("DavidsonLiuSolver called with zero roots."); return -1; }
    if (cpr < 1) { set_err("DavidsonLiuSolver: collapse_per_root_ must be >= 1"); return -1; }
    if (spr < cpr + 1) { set_err("DavidsonLiuSolver: subspace_per_root_ must be >= collapse_per_root_ + 1"); return -1; }
    d->size = size;
    d->nroot = nroot;
    d->collapse_size = (int)((int64_t)cpr * nroot < size ? (int64_t)cpr * nroot : size);
    d->subspace_size = (int)((int64_t)spr * nroot < size ? (int64_t)spr * nroot : size);
    d->discard = 1.0e-7;
    d->orth = 1.0e-12;
    size_t m = (size_t)d->subspace_size;
    d->temp = (double*)calloc(m * (size_t)size, sizeof(double));
    d->b = (double*)calloc(m * (size_t)size, sizeof(double));
    d->r = (double*)calloc(m * (size_t)size, sizeof(double));
    d->sigma = (double*)calloc(m * (size_t)size, sizeof(double));
    d->G = (double*)calloc(m * m, sizeof(double));
    d->alpha = (double*)calloc(m * m, sizeof(double));
    d->lambda = (double*)calloc(m, sizeof(double));
    d->lambda_old = (double*)calloc(m, sizeof(double));
    d->resid = (double*)calloc((size_t)nroot, sizeof(double));
    d->rng = 0x5EEDULL;
    return 0;
}
static void dl_free(dl_t* d) {
    free(d->temp); free(d->b); free(d->r); free(d->sigma);
    free(d->G); free(d->alpha); free(d->lambda); free(d->lambda_old); free(d->resid);
}

/* add_row_and_orthonormalize, davidson_liu_solver.cc:648-687 */
static int dl_add_row(dl_t* d, double* A, int rowsA, const double* B, int rowB) {
    int64_t nc = d->size;
    double* v = A + (size_t)rowsA * nc;
    const double* bb = B + (size_t)rowB * nc;
    memmove(v, bb, (size_t)nc * sizeof(double));
    for (int cycle = 0; cycle < 10; ++cycle) {
        for (int i = 0; i < rowsA; ++i) {
            const double* Ai = A + (size_t)i * nc;
            double dv = ddot(nc, Ai, v);
            for (int64_t I = 0; I < nc; ++I) v[I] -= dv * Ai[I];
        }
        double nrm = sqrt(ddot(nc, v, v));
        if (nrm < d->discard) return 0;
        for (int64_t I = 0; I < nc; ++I) v[I] *= 1. / nrm;
        double mo = 0.0;
        for (int i = 0; i < rowsA; ++i) {
            double o = fabs(ddot(nc, A + (size_t)i * nc, v));
            if (o > mo) mo = o;
        }
        double nn = ddot(nc, v, v);
        if (mo < d->orth && fabs(nn - 1.0) < d->orth) return 1;
    }
    return 0;
}
static int dl_add_rows(dl_t* d, double* A, int rowsA, const double* B, int rowsB) {
    if (rowsA + rowsB > d->subspace_size) { set_err("DavidsonLiuSolver: rowsA + rowsB too large"); return -1; }
    int added = 0;
    for (int j = 0; j < rowsB; ++j)
        if (dl_add_row(d, A, rowsA + added, B, j)) ++added;
    return added;
}
/* project_out_roots, :590-603 (only the first nroot rows) */
static void dl_project_out(dl_t* d, double* v) {
    for (int k = 0; k < d->nroot; ++k) {
        double* vk = v + (size_t)k * d->size;
        for (int j = 0; j < d->nproj; ++j) {
            double ov = 0.0;
            for (int64_t e = d->poff[j]; e < d->poff[j + 1]; ++e) ov += vk[d->pidx[e]] * d->pval[e];
            for (int64_t e = d->poff[j]; e < d->poff[j + 1]; ++e) vk[d->pidx[e]] -= ov * d->pval[e];
        }
    }
}
/* out = alpha^T X over the current basis (temp_->gemm(true,false,...)) */
static void dl_rotate(dl_t* d, const double* X, double* out) {
    int m = d->subspace_size;
    int64_t nc = d->size;
    memset(out, 0, (size_t)m * nc * sizeof(double));
    for (int k = 0; k < m; ++k)
        for (int j = 0; j < d->basis_size; ++j) {
            double a = d->alpha[j * m + k];
            if (a == 0.0) continue;
            daxpy(nc, a, X + (size_t)j * nc, out + (size_t)k * nc);
        }
}

static int dl_solve(dl_t* d, sigma_fn sig, void* user, const double* hdiag, int ng,
                    const int64_t* goff, const int64_t* gidx, const double* gval, int* niter) {
    int m = d->subspace_size;
    int64_t nc = d->size;
    /* setup_guesses, :276-330 */
    if (d->basis_size == 0) {
        if (ng >= d->nroot && ng <= m) {
            memset(d->temp, 0, (size_t)m * nc * sizeof(double));
            for (int g = 0; g < ng; ++g)
                for (int64_t e = goff[g]; e < goff[g + 1]; ++e) d->temp[(size_t)g * nc + gidx[e]] = gval[e];
        } else if (ng == 0) {
            memset(d->temp, 0, (size_t)m * nc * sizeof(double));
            for (int j = 0; j < d->nroot; ++j)
                for (int64_t I = 0; I < nc; ++I) d->temp[(size_t)j * nc + I] = 2.0 * rng_uniform01(&d->rng) - 1.0;
        } else {
            set_err("DavidsonLiuSolver: number of guess vectors must be between 0 and the subspace size");
            return -1;
        }
        int sb = d->nroot > ng ? d->nroot : ng;
        dl_project_out(d, d->temp);
        int added = dl_add_rows(d, d->b, 0, d->temp, sb);
        if (added != sb) { set_err("DavidsonLiuSolver: guess vectors are zero or linearly dependent"); return -1; }
        d->basis_size += added;
    } else if (d->basis_size >= d->nroot) {
        d->basis_size = d->nroot;
        d->sigma_size = 0;
    } else {
        set_err("DavidsonLiuSolver: number of guess vectors must be between 0 and the subspace size");
        return -1;
    }
    if (!sig) { set_err("DavidsonLiuSolver: sigma builder has not been set"); return -1; }
    if (!hdiag) { set_err("DavidsonLiuSolver: h_diag has not been set"); return -1; }

    for (int iter = 0; iter < d->maxiter; ++iter) {
        *niter = iter + 1;
        /* check_orthonormality, :689-730 */
        {
            double thr = d->orth * 3.0, md = 0.0, mo = 0.0;
            for (int i = 0; i < d->basis_size; ++i)
                for (int j = 0; j < d->basis_size; ++j) {
                    double s = ddot(nc, d->b + (size_t)i * nc, d->b + (size_t)j * nc);
                    if (i == j) { if (fabs(s - 1.0) > md) md = fabs(s - 1.0); }
                    else if (fabs(s) > mo) mo = fabs(s);
                }
            if (md > thr) { set_err("DavidsonLiuSolver::check_orthogonality(): eigenvectors are not normalized!"); return -1; }
            if (mo > thr) { set_err("DavidsonLiuSolver::check_orthogonality(): eigenvectors are not orthogonal!"); return -1; }
        }
        /* compute_sigma, :374-383 */
        for (int j = d->sigma_size; j < d->basis_size; ++j) sig(user, d->b + (size_t)j * nc, d->sigma + (size_t)j * nc);
        d->sigma_size = d->basis_size;
        /* form_and_diagonalize_effective_hamiltonian, :385-410 (G = b sigma^T, hermitivitize) */
        int bs = d->basis_size;
        double* Gb = (double*)malloc((size_t)bs * bs * sizeof(double));
        for (int i = 0; i < bs; ++i)
            for (int j = 0; j < bs; ++j) d->G[i * m + j] = ddot(nc, d->b + (size_t)i * nc, d->sigma + (size_t)j * nc);
        for (int i = 0; i < bs; ++i)
            for (int j = 0; j < bs; ++j) Gb[i * bs + j] = 0.5 * (d->G[i * m + j] + d->G[j * m + i]);
        double* ev = (double*)malloc((size_t)bs * sizeof(double));
        double* evec = (double*)malloc((size_t)bs * bs * sizeof(double));
        orc_jacobi(bs, Gb, ev, evec);
        memset(d->alpha, 0, (size_t)m * m * sizeof(double));
        memset(d->lambda, 0, (size_t)m * sizeof(double));
        for (int k = 0; k < bs; ++k) {
            d->lambda[k] = ev[k];
            for (int i = 0; i < bs; ++i) d->alpha[i * m + k] = evec[(size_t)k * bs + i];
        }
        free(Gb); free(ev); free(evec);
        /* form_residual_vectors, :412-430 */
        dl_rotate(d, d->sigma, d->r);
        dl_rotate(d, d->b, d->temp);
        for (int k = 0; k < d->nroot; ++k) {
            double lk = d->lambda[k];
            double* rk = d->r + (size_t)k * nc;
            const double* tk = d->temp + (size_t)k * nc;
            for (int64_t I = 0; I < nc; ++I) rk[I] -= lk * tk[I];
        }
        /* compute_residual_norm, :450-455 */
        for (int k = 0; k < d->nroot; ++k) {
            double* rk = d->r + (size_t)k * nc;
            d->resid[k] = sqrt(ddot(nc, rk, rk));
        }
        /* form_correction_vectors, :432-448 */
        for (int k = 0; k < d->nroot; ++k) {
            double* rk = d->r + (size_t)k * nc;
            double lk = d->lambda[k];
            for (int64_t I = 0; I < nc; ++I) {
                double den = lk - hdiag[I];
                if (fabs(den) > 1.0e-6) rk[I] /= den; else rk[I] = 0.0;
            }
        }
        dl_project_out(d, d->r);
        /* normalize_vectors, :457-465 */
        for (int k = 0; k < d->nroot; ++k) {
            double* rk = d->r + (size_t)k * nc;
            double nr = sqrt(ddot(nc, rk, rk));
            for (int64_t I = 0; I < nc; ++I) rk[I] /= nr;
        }
        /* check_convergence, :467-486 */
        int ne = 0, nr = 0;
        for (int k = 0; k < d->nroot; ++k) {
            ne += fabs(d->lambda[k] - d->lambda_old[k]) < d->e_conv;
            nr += d->resid[k] < d->r_conv;
        }
        int e_ok = ne == d->nroot, r_ok = nr == d->nroot;
        if ((e_ok && r_ok) || d->basis_size == nc) goto results_true;
        /* subspace_collapse, :538-588 */
        if (d->basis_size + d->nroot > m) {
            int cs = d->collapse_size > d->nroot ? d->collapse_size : d->nroot;
            if (cs > d->basis_size) cs = d->basis_size;
            dl_rotate(d, d->b, d->temp);
            memset(d->b, 0, (size_t)m * nc * sizeof(double));
            int added = dl_add_rows(d, d->b, 0, d->temp, cs);
            dl_rotate(d, d->sigma, d->temp);
            memcpy(d->sigma, d->temp, (size_t)m * nc * sizeof(double));
            if (added != cs) { set_err("DavidsonLiuSolver: collapse_vectors generated less vectors than expected"); return -1; }
            d->basis_size = cs;
            d->sigma_size = cs;
        }
        int num_to_add = d->nroot < m - d->basis_size ? d->nroot : m - d->basis_size;
        int added = dl_add_rows(d, d->b, d->basis_size, d->r, num_to_add);
        if (added < 0) return -1;
        d->basis_size += added;
        int missing = num_to_add - added;
        if (missing > 0) {
            /* add_random_vectors fills temp_ then project_out_roots(temp_) in the reference */
            memset(d->temp, 0, (size_t)m * nc * sizeof(double));
            for (int j = 0; j < missing; ++j)
                for (int64_t I = 0; I < nc; ++I) d->temp[(size_t)j * nc + I] = 2.0 * rng_uniform01(&d->rng) - 1.0;
            dl_project_out(d, d->temp);
            int ra = dl_add_rows(d, d->b, d->basis_size, d->temp, missing);
            d->basis_size += ra;
            added += ra;
        }
        if (added == 0) {
            if (r_ok) goto results_true;
            goto results_false;
        }
        memcpy(d->lambda_old, d->lambda, (size_t)m * sizeof(double));
    }
results_false:
    {
        /* get_results also on the non-converged path, :488-501 */
        memcpy(d->lambda_old, d->lambda, (size_t)m * sizeof(double));
        dl_rotate(d, d->b, d->temp);
        memset(d->b, 0, (size_t)m * nc * sizeof(double));
        if (dl_add_rows(d, d->b, 0, d->temp, d->nroot) != d->nroot) { set_err("DavidsonLiuSolver: get_results generated less vectors than expected"); return -1; }
        return 0;
    }
results_true:
    memcpy(d->lambda_old, d->lambda, (size_t)m * sizeof(double));
    dl_rotate(d, d->b, d->temp);
    memset(d->b, 0, (size_t)m * nc * sizeof(double));
    if (dl_add_rows(d, d->b, 0, d->temp, d->nroot) != d->nroot) { set_err("DavidsonLiuSolver: get_results generated less vectors than expected"); return -1; }
    return 1;
}

typedef struct {
    const double* M;
    int64_t n;
} dense_user;
static void dense_sigma(void* u, const double* b, double* s) {
    const dense_user* du = (const dense_user*)u;
    for (int64_t i = 0; i < du->n; ++i) {
        double r = 0.0;
        for (int64_t j = 0; j < du->n; ++j) r += du->M[i * du->n + j] * b[j];
        s[i] = r;
    }
}

int orc_dl_dense(int64_t size, int nroot, const orc_dl_opts* o, const double* M,
                 const double* hdiag, int ng, const int64_t* goff, const int64_t* gidx,
                 const double* gval, int np, const int64_t* poff, const int64_t* pidx,
                 const double* pval, double* evals, double* evecs, int* niter) {
    dl_t d;
    if (dl_init(&d, size, nroot, o->collapse_per_root, o->subspace_per_root) < 0) return -1;
    d.maxiter = o->maxiter;
    d.e_conv = o->e_convergence;
    d.r_conv = o->r_convergence;
    d.nproj = np;
    d.poff = poff;
    d.pidx = pidx;
    d.pval = pval;
    dense_user u = {M, size};
    int it = 0;
    int rc = dl_solve(&d, M ? dense_sigma : NULL, &u, hdiag, ng, goff, gidx, gval, &it);
    if (rc >= 0) {
        for (int k = 0; k < nroot; ++k) evals[k] = d.lambda[k];
        if (evecs) memcpy(evecs, d.b, (size_t)nroot * size * sizeof(double));
    }
    if (niter) *niter = it;
    dl_free(&d);
    return rc;
}

/* ---------------------------------------------------------------------------------------- */
/* Initial guess.                                                                            */
/* spin2(lhs, rhs), sparse_ci/determinant.hpp:910-955 */
static double spin2_el(int n, det2 I, det2 J) {
    if (popc(I.a) != popc(J.a) || popc(I.b) != popc(J.b)) return 0.0;
    int nadiff = popc(I.a ^ J.a) / 2, nbdiff = popc(I.b ^ J.b) / 2;
    int na = popc(I.a), nb = popc(I.b), npair = popc(I.a & I.b);
    double Ms = 0.5 * (double)(na - nb), me = 0.0;
    if (nadiff == 0 && nbdiff == 0) me += Ms * (Ms + 1.0) + (double)nb - (double)npair;
    if (nadiff == 1 && nbdiff == 1) {
        int i = -1, j = -1;
        for (int p = 0; p < n; ++p) {
            if (((J.a >> p) & 1) && ((I.b >> p) & 1) && !((J.b >> p) & 1) && !((I.a >> p) & 1)) i = p;
            if (((J.b >> p) & 1) && ((I.a >> p) & 1) && !((J.a >> p) & 1) && !((I.b >> p) & 1)) j = p;
        }
        if (i != j && i >= 0 && j >= 0) {
            double sign = sign_below(J.a, i) * sign_below(J.b, j) * sign_below(I.a, j) * sign_below(I.b, i);
            me -= sign;
        }
    }
    return me;
}

/* find_integer_groups, helpers/helpers.cc:237-262 */
static int near_int(double v, double t) { return fabs(v - round(v)) <= t; }
static int integer_groups(const double* v, int n, double toll, int* gm, int* gs, int* ge) {
    int ng = 0, it = 0;
    while (it != n) {
        while (it < n && !near_int(v[it], toll)) ++it;
        if (it != n) {
            int start = it, j = it;
            while (j + 1 < n && !(near_int(v[j], toll) && fabs(v[j] - v[j + 1]) > 2 * toll)) ++j;
            int end = j; /* adjacent_find result, or last element */
            gm[ng] = (int)llround(v[start]);
            gs[ng] = start;
            ge[ng] = end + 1;
            ++ng;
            it = end + 1;
        }
    }
    return ng;
}

typedef struct {
    det2* d;
    int n, cap;
} detvec;
static void dv_push(detvec* v, det2 x) {
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 64;
        v->d = (det2*)realloc(v->d, (size_t)v->cap * sizeof(det2));
    }
    v->d[v->n++] = x;
}
static int dv_has(const detvec* v, det2 x) {
    for (int i = 0; i < v->n; ++i)
        if (v->d[i].a == x.a && v->d[i].b == x.b) return 1;
    return 0;
}
/* enforce_spin_completeness, sparse_ci/determinant_functions.cc:133-194 */
static void spin_complete(detvec* dets, int nmo) {
    int det_size = dets->n;
    for (int I = 0; I < det_size; ++I) {
        det2 det = dets->d[I];
        int open[64], closed[64], naopen = 0, nbopen = 0, nclosed = 0;
        for (int i = 0; i < nmo; ++i) {
            int a = (det.a >> i) & 1, b = (det.b >> i) & 1;
            if (a && !b) { open[naopen + nbopen] = i; naopen++; }
            else if (!a && b) { open[naopen + nbopen] = i; nbopen++; }
            else if (a && b) closed[nclosed++] = i;
        }
        if (naopen + nbopen == 0) continue;
        int8_t bits[64];
        for (int i = 0; i < nbopen; ++i) bits[i] = 0;
        for (int i = nbopen; i < naopen + nbopen; ++i) bits[i] = 1;
        do {
            det2 nd = {0, 0};
            for (int c = 0; c < nclosed; ++c) { nd.a |= 1ULL << closed[c]; nd.b |= 1ULL << closed[c]; }
            for (int o = 0; o < naopen + nbopen; ++o) {
                if (bits[o]) nd.a |= 1ULL << open[o]; else nd.b |= 1ULL << open[o];
            }
            if (!dv_has(dets, nd)) dv_push(dets, nd);
        } while (next_perm(bits, naopen + nbopen));
    }
}

typedef struct {
    orc_ctx* ctx;
} fci_user;
static void fci_sigma(void* u, const double* b, double* s) {
    orc_ctx* ctx = ((fci_user*)u)->ctx;
    orc_sigma(ctx, b, s, ORC_ALL);
}

int orc_fci_solve(orc_ctx* ctx, const orc_fci_opts* o, double* energies, double* spin2,
                  double* evecs, int* niter) {
    int64_t nd = orc_ndet(ctx), nbs = ctx->sp[1].nstr;
    int nroot = o->nroot;
    dl_t d;
    if (dl_init(&d, nd, nroot, o->collapse_per_root, o->subspace_per_root) < 0) return -1;
    d.maxiter = o->maxiter;
    d.e_conv = o->e_convergence;
    d.r_conv = o->r_convergence;
    double* hdiag = (double*)malloc((size_t)nd * sizeof(double));
    orc_hdiag(ctx, hdiag);
    int64_t ngs = (int64_t)o->guess_per_root * nroot;
    if (ngs > nd) ngs = nd; /* fci_solver.cc:217 */
    int num_guess_states = (int)ngs;
    /* initial_guess_generate_dets, fci_solver_initial_guess.cc:44-86 */
    int64_t ngd = (int64_t)num_guess_states * o->ndets_per_guess;
    if (ngd > nd) ngd = nd;
    double* ve = (double*)malloc((size_t)ngd * sizeof(double));
    int64_t* vi = (int64_t*)malloc((size_t)ngd * sizeof(int64_t));
    for (int64_t i = 0; i < ngd; ++i) { ve[i] = 1.7976931348623157e308; vi[i] = 0; }
    double emax = 1.7976931348623157e308;
    int64_t added = 0;
    for (int64_t I = 0; I < nd; ++I) {
        double e = hdiag[I];
        if (e < emax || added < num_guess_states) {
            /* pop_back, then insert before the first element with e < its energy */
            int64_t len = ngd - 1, pos = len;
            for (int64_t t = 0; t < len; ++t) if (e < ve[t]) { pos = t; break; }
            for (int64_t t = len; t > pos; --t) { ve[t] = ve[t - 1]; vi[t] = vi[t - 1]; }
            ve[pos] = e;
            vi[pos] = I;
            emax = ve[ngd - 1];
            added++;
        }
    }
    detvec gd = {0};
    for (int64_t i = 0; i < ngd; ++i) {
        det2 x = {ctx->sp[0].str[vi[i] / nbs], ctx->sp[1].str[vi[i] % nbs]};
        dv_push(&gd, x);
    }
    free(ve);
    free(vi);
    spin_complete(&gd, ctx->n);
    int G = gd.n;
    int64_t* gpos = (int64_t*)malloc((size_t)G * sizeof(int64_t));
    for (int i = 0; i < G; ++i) gpos[i] = det_index(ctx, gd.d[i]);
    /* find_initial_guess_det, sparse_initial_guess.cc:71-264 */
    double* S2 = (double*)malloc((size_t)G * G * sizeof(double));
    double* H = (double*)malloc((size_t)G * G * sizeof(double));
    for (int I = 0; I < G; ++I)
        for (int J = I; J < G; ++J) {
            double s = spin2_el(ctx->n, gd.d[I], gd.d[J]);
            S2[I * G + J] = S2[J * G + I] = s;
            double h = slater_rules(ctx, gd.d[I].a, gd.d[I].b, gd.d[J].a, gd.d[J].b);
            H[I * G + J] = H[J * G + I] = h;
        }
    double* s2val = (double*)malloc((size_t)G * sizeof(double));
    double* s2vec = (double*)malloc((size_t)G * G * sizeof(double)); /* column k at [k*G] */
    orc_jacobi(G, S2, s2val, s2vec);
    /* HS2 = U^T H U */
    double* HU = (double*)calloc((size_t)G * G, sizeof(double));
    double* HS2 = (double*)calloc((size_t)G * G, sizeof(double));
    for (int i = 0; i < G; ++i)
        for (int k = 0; k < G; ++k) {
            double s = 0.0;
            for (int j = 0; j < G; ++j) s += H[i * G + j] * s2vec[(size_t)k * G + j];
            HU[i * G + k] = s;
        }
    for (int a = 0; a < G; ++a)
        for (int k = 0; k < G; ++k) {
            double s = 0.0;
            for (int i = 0; i < G; ++i) s += s2vec[(size_t)a * G + i] * HU[i * G + k];
            HS2[a * G + k] = s;
        }
    double* mult = (double*)malloc((size_t)G * sizeof(double));
    for (int i = 0; i < G; ++i) mult[i] = sqrt(1.0 + 4.0 * s2val[i]);
    int *gm = (int*)malloc(sizeof(int) * (size_t)G), *gs = (int*)malloc(sizeof(int) * (size_t)G),
        *ge = (int*)malloc(sizeof(int) * (size_t)G);
    int ngroups = integer_groups(mult, G, 1.0e-6, gm, gs, ge);
    /* per group: energies, C block = U[:,block] * evecs(HS2 block) */
    double** genergy = (double**)calloc((size_t)ngroups, sizeof(double*));
    double** gC = (double**)calloc((size_t)ngroups, sizeof(double*));
    int target = -1;
    for (int g = 0; g < ngroups; ++g) {
        int bsz = ge[g] - gs[g];
        double* blk = (double*)malloc((size_t)bsz * bsz * sizeof(double));
        for (int i = 0; i < bsz; ++i)
            for (int j = 0; j < bsz; ++j) blk[i * bsz + j] = HS2[(gs[g] + i) * G + gs[g] + j];
        double* bev = (double*)malloc((size_t)bsz * sizeof(double));
        double* bvec = (double*)malloc((size_t)bsz * bsz * sizeof(double));
        orc_jacobi(bsz, blk, bev, bvec);
        genergy[g] = bev;
        for (int i = 0; i < bsz; ++i) bev[i] += ctx->e_core;
        gC[g] = (double*)calloc((size_t)G * bsz, sizeof(double)); /* [I*bsz + r] */
        for (int I = 0; I < G; ++I)
            for (int r = 0; r < bsz; ++r) {
                double s = 0.0;
                for (int j = 0; j < bsz; ++j) s += s2vec[(size_t)(gs[g] + j) * G + I] * bvec[(size_t)r * bsz + j];
                gC[g][I * bsz + r] = s;
            }
        free(blk);
        free(bvec);
        if (gm[g] == o->multiplicity) target = g; /* std::map keyed by multiplicity: last wins */
    }
    int rc = 0;
    int64_t *goff = NULL, *gidx = NULL, *poff = NULL, *pidx = NULL;
    double *gval = NULL, *pval = NULL;
    if (target < 0) {
        set_err("No guess with the requested multiplicity was found.");
        rc = -1;
        goto cleanup;
    }
    {
        int bsz = ge[target] - gs[target];
        if (bsz < num_guess_states) {
            set_err("Found fewer guesses with the requested multiplicity than were requested. Increase the value of DL_DETS_PER_GUESS");
            rc = -1;
            goto cleanup;
        }
        goff = (int64_t*)malloc(sizeof(int64_t) * (size_t)(num_guess_states + 1));
        gidx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(num_guess_states * G));
        gval = (double*)malloc(sizeof(double) * (size_t)(num_guess_states * G));
        double gmax = -1.7976931348623157e308;
        for (int r = 0; r < num_guess_states; ++r) {
            goff[r] = (int64_t)r * G;
            for (int I = 0; I < G; ++I) {
                gidx[r * G + I] = gpos[I];
                gval[r * G + I] = gC[target][I * bsz + r];
            }
            if (genergy[target][r] > gmax) gmax = genergy[target][r];
        }
        goff[num_guess_states] = (int64_t)num_guess_states * G;
        /* bad roots: other multiplicities below the highest selected guess energy.
         * std::map iterates multiplicities in ascending order. */
        int np = 0;
        poff = (int64_t*)malloc(sizeof(int64_t) * (size_t)(G + 1));
        pidx = (int64_t*)malloc(sizeof(int64_t) * (size_t)G * G);
        pval = (double*)malloc(sizeof(double) * (size_t)G * G);
        poff[0] = 0;
        int* order = (int*)malloc(sizeof(int) * (size_t)ngroups);
        int no = 0;
        for (int g = 0; g < ngroups; ++g) {
            /* std::map keeps one entry per multiplicity (the last group assigned) */
            int keep = 1;
            for (int h = g + 1; h < ngroups; ++h) if (gm[h] == gm[g]) keep = 0;
            if (keep) order[no++] = g;
        }
        for (int i = 0; i < no; ++i)
            for (int j = i + 1; j < no; ++j)
                if (gm[order[j]] < gm[order[i]]) { int t = order[i]; order[i] = order[j]; order[j] = t; }
        for (int oi = 0; oi < no; ++oi) {
            int g = order[oi];
            if (gm[g] == o->multiplicity) continue;
            int bsz2 = ge[g] - gs[g];
            for (int r = 0; r < bsz2; ++r)
                if (genergy[g][r] < gmax) {
                    for (int I = 0; I < G; ++I) {
                        pidx[poff[np] + I] = gpos[I];
                        pval[poff[np] + I] = gC[g][I * bsz2 + r];
                    }
                    poff[np + 1] = poff[np] + G;
                    np++;
                }
        }
        free(order);
        d.nproj = np;
        d.poff = poff;
        d.pidx = pidx;
        d.pval = pval;
        fci_user u = {ctx};
        int it = 0;
        rc = dl_solve(&d, fci_sigma, &u, hdiag, num_guess_states, goff, gidx, gval, &it);
        if (niter) *niter = it;
        if (rc >= 0) {
            for (int r = 0; r < nroot; ++r) {
                energies[r] = d.lambda[r];
                if (spin2) spin2[r] = orc_spin2(ctx, d.b + (size_t)r * nd);
            }
            if (evecs) memcpy(evecs, d.b, (size_t)nroot * nd * sizeof(double));
        }
    }
cleanup:
    for (int g = 0; g < ngroups; ++g) { free(genergy[g]); free(gC[g]); }
    free(genergy); free(gC); free(gm); free(gs); free(ge); free(mult);
    free(HU); free(HS2); free(s2val); free(s2vec); free(S2); free(H); free(gpos);
    free(gd.d); free(hdiag); free(goff); free(gidx); free(gval); free(poff); free(pidx); free(pval);
    dl_free(&d);
    return rc;
}

/* ---------------------------------------------------------------------------------------- */
/* Seeded integral generator (the synthetic inputs of the benchmark configurations):
 *   RNG: SplitMix64 seeded with `seed`; uniform u = ((x >> 11) + 0.5) * 2^-53 in (0,1);
 *   normal draw = sqrt(-2 ln u1) * cos(2 pi u2) from two consecutive uniforms (std dev 1).
 *   h_pp = -2.0 + 1.5 p / (n-1) (n > 1; h_00 = -2.0 when n == 1);
 *   h_pq = h_qp = 0.05 N for p < q, drawn in row-major order (p outer, q inner);
 *   for K = 0..2n-1: L^K_pq = L^K_qp = 0.1 N for p <= q (p outer, q inner), then L^K_pp += 0.5;
 *   (pq|rs) = sum_K L^K_pq L^K_rs, summed in ascending K.                                       */
static double rng_normal(uint64_t* s) {
    double u1 = rng_uniform01(s), u2 = rng_uniform01(s);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925286766559 * u2);
}
void orc_seeded_integrals(int n, uint64_t seed, double* h, double* eri) {
    uint64_t st = seed;
    for (int p = 0; p < n; ++p) h[p * n + p] = (n > 1) ? -2.0 + 1.5 * (double)p / (double)(n - 1) : -2.0;
    for (int p = 0; p < n; ++p)
        for (int q = p + 1; q < n; ++q) {
            double v = 0.05 * rng_normal(&st);
            h[p * n + q] = v;
            h[q * n + p] = v;
        }
    int nk = 2 * n;
    size_t n2 = (size_t)n * n;
    double* L = (double*)malloc((size_t)nk * n2 * sizeof(double));
    for (int K = 0; K < nk; ++K) {
        double* l = L + (size_t)K * n2;
        for (int p = 0; p < n; ++p)
            for (int q = p; q < n; ++q) {
                double v = 0.1 * rng_normal(&st);
                l[p * n + q] = v;
                l[q * n + p] = v;
            }
        for (int p = 0; p < n; ++p) l[p * n + p] += 0.5;
    }
    for (size_t pq = 0; pq < n2; ++pq)
        for (size_t rs = 0; rs < n2; ++rs) {
            double s = 0.0;
            for (int K = 0; K < nk; ++K) s += L[(size_t)K * n2 + pq] * L[(size_t)K * n2 + rs];
            eri[pq * n2 + rs] = s;
        }
    free(L);
}

/* make_active_space_ints, integrals/active_space_integrals.cc:578-583 */
void orc_chem_to_forte(int n, const double* eri, double* tei_aa, double* tei_ab) {
    for (int p = 0; p < n; ++p)
        for (int q = 0; q < n; ++q)
            for (int r = 0; r < n; ++r)
                for (int s = 0; s < n; ++s) {
                    double prqs = eri[IDX4(p, r, q, s, n)], psqr = eri[IDX4(p, s, q, r, n)];
                    tei_ab[IDX4(p, q, r, s, n)] = prqs;
                    tei_aa[IDX4(p, q, r, s, n)] = prqs - psqr;
                }
}
