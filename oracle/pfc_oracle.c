/*
 * pfc_oracle.c -- CPU restatement of PFCSieve's sieve path, used ONLY as a
 * test checker (tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg).  Nothing in the product library links or calls this file.
 *
 * It restates, in plain C with unsigned __int128, what the reference computes
 * on its OpenCL device and in its host driver, following the reference files
 * cited at each function (paths relative to the PFCSieve source tree):
 *
 *   Montgomery arithmetic, R = 2^64 ...... kernels/iterate.cl:44-57,
 *                                          verifyprime.cpp:23-59
 *   base-2 strong PRP .................... kernels/getsegprimes.cl:79-108
 *   candidate set (wheel-30, sieve 7..113, PRP; primes < 114)
 *                                          kernels/getsegprimes.cl:170-232,
 *                                          kernels/addsmallprimes.cl:21-40,
 *                                          cl_sieve.cpp:1598-1615
 *   wheel offset ......................... cl_sieve.cpp:369-430
 *   tables ............................... cl_sieve.cpp:824-1326
 *   start residues ....................... kernels/setup.cl:49-288
 *   residue walk ......................... kernels/iterate.cl:59-198,
 *                                          cl_sieve.cpp:1697-1719
 *   checksum terms ....................... kernels/check.cl:45-204
 *   harvest (sort, isPrime filter, n+c) .. cl_sieve.cpp:433-593
 *   isPrime / verify ..................... verifyprime.cpp:109-218
 *
 * The reference's prime lists come from primesieve 12.3 (a prebuilt third-party
 * library; primesieve_generate_primes(lo, hi) returns the primes in [lo, hi]).
 * Prime lists are mathematically unique, so a plain sieve of Eratosthenes
 * stands in for it here (orc_primes).
 *
 * Pinning: every function here is checked in tests/test_oracle_*.py against
 * golden vectors the reference holds (the 16 self-test triples of
 * cl_sieve.cpp:1787-2080, the p = 2^64-59 constants of verify.cl:46, the
 * 34!, 101#, 44!/44# constants of verifyprime.cpp:155,172,191), against the
 * reference's own verifyprime.cpp/putil.c built into oracle/_ref, and against
 * independent Python big-integer brute force on small ranges.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;

enum { ORC_F = 1u, ORC_P = 2u, ORC_C = 4u };
enum { ORC_FACTORIAL = 0, ORC_PRIMORIAL = 1, ORC_COMPOSITORIAL = 2 };

typedef struct { uint64_t p; int32_t nc; int32_t type; } orc_factor;  /* cl_sieve.h:8-12 */

/* ---------------------------------------------------------------- Montgomery */

/* Newton iteration for p^-1 mod 2^64 (verifyprime.cpp:23-28). */
uint64_t orc_invert(uint64_t p) {
    uint64_t inv = 1, prev = 0;
    while (inv != prev) { prev = inv; inv *= 2 - p * inv; }
    return inv;
}

/* a*b*2^-64 mod p, canonical for a*b < p*2^64 (iterate.cl:44-50). */
uint64_t orc_mmul(uint64_t a, uint64_t b, uint64_t p, uint64_t q) {
    u128 t = (u128)a * b;
    uint64_t lo = (uint64_t)t, hi = (uint64_t)(t >> 64);
    uint64_t m = lo * q;
    uint64_t mp = (uint64_t)(((u128)m * p) >> 64);
    uint64_t r = hi - mp;
    return hi < mp ? r + p : r;
}

/* a+b mod p for a,b < p, exact across the 2^64 wrap (iterate.cl:52-57). */
uint64_t orc_madd(uint64_t a, uint64_t b, uint64_t p) {
    uint64_t c = (a >= p - b) ? p : 0;
    return a + b - c;
}

typedef struct { uint64_t p, q, one, two, nmo, r2; } orc_mont;

/* Per-prime constants as getsegprimes.cl:214-217 and setup.cl:68-73 build them. */
static orc_mont mont_init(uint64_t p) {
    orc_mont m;
    m.p = p;
    m.q = orc_invert(p);
    m.one = (0 - p) % p;
    m.nmo = p - m.one;
    m.two = orc_madd(m.one, m.one, p);
    m.r2 = orc_madd(m.two, m.two, p);
    for (int i = 0; i < 5; ++i) m.r2 = orc_mmul(m.r2, m.r2, p, m.q);  /* (4R)^(2^5)/R^31 = R^2 */
    return m;
}

/* ---------------------------------------------------------------- primality */

/* Base-2 strong probable-prime test in Montgomery form (getsegprimes.cl:79-108). */
int orc_strong_prp2(uint64_t N) {
    if (N < 3 || !(N & 1)) return 0;
    orc_mont m = mont_init(N);
    int t = __builtin_ctzll(N - 1);
    uint64_t d = N >> t;
    /* OpenCL masks shift counts to 6 bits, so d == 1 (N = 2^t+1) shifts by 0. */
    uint64_t bit = 0x8000000000000000ull >> ((__builtin_clzll(d) + 1) & 63);
    uint64_t a = m.two;
    while (bit) {
        a = orc_mmul(a, a, N, m.q);
        if (d & bit) a = orc_madd(a, a, N);
        bit >>= 1;
    }
    if (a == m.one || a == m.nmo) return 1;
    for (int s = 1; s < t; ++s) {
        a = orc_mmul(a, a, N, m.q);
        if (a == m.nmo) return 1;
    }
    return 0;
}

/* Strong PRP to an arbitrary base (verifyprime.cpp:66-104). */
static int strong_prp_base(uint32_t base, const orc_mont* m, int t, uint64_t d, uint64_t bit) {
    uint64_t a = orc_mmul(base, m->r2, m->p, m->q);
    const uint64_t mb = a;
    while (bit) {
        a = orc_mmul(a, a, m->p, m->q);
        if (d & bit) a = orc_mmul(a, mb, m->p, m->q);
        bit >>= 1;
    }
    if (a == m->one || a == m->nmo) return 1;
    for (int s = 1; s < t; ++s) {
        a = orc_mmul(a, a, m->p, m->q);
        if (a == m->nmo) return 1;
    }
    return 0;
}

/* Deterministic 7-base Miller-Rabin for 64-bit p (verifyprime.cpp:109-145). */
int orc_is_prime(uint64_t p) {
    static const uint32_t bases[7] = {2, 325, 9375, 28178, 450775, 9780504, 1795265022};
    if (p % 2 == 0) return 0;
    orc_mont m = mont_init(p);
    int t = __builtin_ctzll(p - 1);
    uint64_t d = p >> t;
    /* x86 masks 64-bit shift counts to 6 bits: d == 1 shifts by 0, as the
     * reference binary does. */
    uint64_t bit = 0x8000000000000000ull >> ((__builtin_clzll(d) + 1) & 63);
    for (int i = 0; i < 7; ++i) {
        uint32_t b = bases[i];
        if (b >= p) { b = (uint32_t)(b % p); if (b == 0) continue; }
        if (!strong_prp_base(b, &m, t, d, bit)) return 0;
    }
    return 1;
}

/* ---------------------------------------------------------------- prime lists */

/* Primes in [lo, hi] (inclusive, hi < 2^32): the semantics of
 * primesieve_generate_primes(lo, hi) at cl_sieve.cpp:850,996,1000,1061,1158,
 * 1472,1476,1498.  Returns the count; writes at most cap values. */
size_t orc_primes(uint32_t lo, uint32_t hi, uint32_t* out, size_t cap) {
    if (hi < 2 || lo > hi) return 0;
    uint64_t H = hi;
    uint32_t root = 1;
    while ((uint64_t)(root + 1) * (root + 1) <= H) ++root;
    uint8_t* small = (uint8_t*)calloc(root + 1, 1);
    for (uint32_t i = 2; (uint64_t)i * i <= root; ++i)
        if (!small[i]) for (uint32_t j = i * i; j <= root; j += i) small[j] = 1;
    size_t n = 0;
    const uint64_t SEG = 1u << 20;
    uint8_t* seg = (uint8_t*)malloc(SEG);
    uint64_t L = lo < 2 ? 2 : lo;
    for (uint64_t s = L; s <= H; s += SEG) {
        uint64_t e = s + SEG - 1;
        if (e > H) e = H;
        memset(seg, 0, (size_t)(e - s + 1));
        for (uint32_t i = 2; i <= root; ++i) {
            if (small[i]) continue;
            uint64_t st = (uint64_t)i * i;
            if (st < s) st = ((s + i - 1) / i) * (uint64_t)i;
            for (uint64_t j = st; j <= e; j += i) seg[j - s] = 1;
        }
        for (uint64_t x = s; x <= e; ++x)
            if (!seg[x - s]) { if (n < cap) out[n] = (uint32_t)x; ++n; }
    }
    free(seg);
    free(small);
    return n;
}

/* ---------------------------------------------------------------- candidates */

/* First N >= start with gcd(N, 30) = 1 and its mod-30 wheel index
 * (cl_sieve.cpp:369-430).  Index k means the gaps that follow N are
 * {4,2,4,2,4,6,2,6}[k], [k+1], ... (cyclic); N = 1, 7, 11, 13, 17, 19, 23, 29
 * (mod 30) gives k = 7, 0, 1, 2, 3, 4, 5, 6. */
void orc_wheel_offset(uint64_t* start, int32_t* idx) {
    static const uint8_t res[8] = {1, 7, 11, 13, 17, 19, 23, 29};
    static const int32_t kidx[8] = {7, 0, 1, 2, 3, 4, 5, 6};
    uint64_t N = *start;
    for (;;) {
        uint32_t r = (uint32_t)(N % 30);
        for (int i = 0; i < 8; ++i)
            if (r == res[i]) { *start = N; *idx = kidx[i]; return; }
        ++N;
    }
}

static const uint32_t sieve_primes[27] = {7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59,
                                          61, 67, 71, 73, 79, 83, 89, 97, 101, 103, 107, 109, 113};
static const uint32_t low_primes[29] = {3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53,
                                        59, 61, 67, 71, 73, 79, 83, 89, 97, 101, 103, 107, 109, 113};

static int survives_wheel_sieve(uint64_t c) {
    if (!(c & 1) || c % 3 == 0 || c % 5 == 0) return 0;
    for (int i = 0; i < 27; ++i)
        if (c % sieve_primes[i] == 0) return 0;
    return 1;
}

/* The reference's candidate set for [pmin, pmax): primes below 114 from
 * addsmallprimes (addsmallprimes.cl:25-31, host clamp cl_sieve.cpp:1599-1605),
 * then every wheel-30 number in [max(pmin,114), pmax) with no factor 7..113
 * that passes the base-2 strong PRP (getsegprimes.cl:193-223).  Returns the
 * count, writes at most cap values in ascending order. */
size_t orc_candidates(uint64_t pmin, uint64_t pmax, uint64_t* out, size_t cap) {
    size_t n = 0;
    for (int i = 0; i < 29; ++i) {
        uint64_t p = low_primes[i];
        if (p >= pmin && p < pmax) { if (n < cap) out[n] = p; ++n; }
    }
    uint64_t lo = pmin < 114 ? 114 : pmin;
    if (lo >= pmax) return n;
    /* Blocks processed in parallel; results concatenated in order. */
    const uint64_t BLK = 1u << 16;
    uint64_t span = pmax - lo;
    uint64_t nblk = span / BLK + 1;
    size_t* counts = (size_t*)calloc(nblk, sizeof(size_t));
    uint64_t** bufs = (uint64_t**)calloc(nblk, sizeof(uint64_t*));
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t b = 0; b < (int64_t)nblk; ++b) {
        uint64_t s = lo + (uint64_t)b * BLK;
        u128 e128 = (u128)s + BLK;
        uint64_t e = e128 > pmax ? pmax : (uint64_t)e128;
        if (s >= e) continue;
        uint64_t* buf = (uint64_t*)malloc(sizeof(uint64_t) * (BLK / 4 + 16));
        size_t k = 0;
        for (u128 c = s | 1; c < e; c += 2)
            if (survives_wheel_sieve((uint64_t)c) && orc_strong_prp2((uint64_t)c)) buf[k++] = (uint64_t)c;
        bufs[b] = buf;
        counts[b] = k;
    }
    for (uint64_t b = 0; b < nblk; ++b) {
        for (size_t k = 0; k < counts[b]; ++k) { if (n < cap) out[n] = bufs[b][k]; ++n; }
        free(bufs[b]);
    }
    free(bufs);
    free(counts);
    return n;
}

/* ---------------------------------------------------------------- tables */

/* Legendre exponent of prime r in N! and the bit below its leading one
 * (cl_sieve.cpp:824-839). */
static void get_power(uint32_t r, uint32_t N, uint32_t* e, uint32_t* bit) {
    uint32_t tot = 0;
    uint64_t pk = r;
    while (pk <= N) { tot += (uint32_t)(N / pk); pk *= r; }
    uint32_t b = 0x80000000u;
    if (tot > 1) b >>= (__builtin_clz(tot) + 1);
    *e = tot;
    *bit = b;
}

/* Factorial power table for N = nmin-1 (cl_sieve.cpp:842-885): entry 0 is the
 * prime 2 alone; later entries multiply runs of consecutive primes with equal
 * exponent while the product stays below 2^64. */
size_t orc_power_table(uint32_t N, uint64_t* prod, uint32_t* exps, uint32_t* bits, size_t cap) {
    size_t np = orc_primes(2, N, NULL, 0);
    uint32_t* pr = (uint32_t*)malloc(sizeof(uint32_t) * (np + 1));
    orc_primes(2, N, pr, np);
    uint32_t* pe = (uint32_t*)malloc(sizeof(uint32_t) * (np + 1));
    uint32_t* pb = (uint32_t*)malloc(sizeof(uint32_t) * (np + 1));
    for (size_t i = 0; i < np; ++i) get_power(pr[i], N, &pe[i], &pb[i]);
    size_t m = 0;
    if (np) {
        if (cap) { prod[0] = pr[0]; exps[0] = pe[0]; bits[0] = pb[0]; }
        m = 1;
        for (size_t i = 1; i < np; ++m) {
            uint64_t acc = pr[i];
            uint32_t e = pe[i], b = pb[i];
            for (++i; i < np && pe[i] == e; ++i) {
                u128 pp = (u128)acc * pr[i];
                if (pp > 0xFFFFFFFFFFFFFFFFull) break;
                acc = (uint64_t)pp;
            }
            if (m < cap) { prod[m] = acc; exps[m] = e; bits[m] = b; }
        }
    }
    free(pr); free(pe); free(pb);
    return m;
}

/* Primorial product table: primes <= N multiplied greedily below 2^64
 * (cl_sieve.cpp:987-1020). */
size_t orc_prime_products(uint32_t N, uint64_t* prod, size_t cap) {
    size_t np = orc_primes(2, N, NULL, 0);
    uint32_t* pr = (uint32_t*)malloc(sizeof(uint32_t) * (np + 1));
    orc_primes(2, N, pr, np);
    size_t m = 0;
    for (size_t i = 0; i < np; ++m) {
        uint64_t acc = pr[i];
        for (++i; i < np; ++i) {
            u128 pp = (u128)acc * pr[i];
            if (pp > 0xFFFFFFFFFFFFFFFFull) break;
            acc = (uint64_t)pp;
        }
        if (m < cap) prod[m] = acc;
    }
    free(pr);
    return m;
}

/* Compositorial product table: composites in [4, N] multiplied greedily below
 * 2^64 (cl_sieve.cpp:1151-1193). */
size_t orc_comp_products(uint32_t N, uint64_t* prod, size_t cap) {
    size_t np = orc_primes(2, N, NULL, 0);
    uint32_t* pr = (uint32_t*)malloc(sizeof(uint32_t) * (np + 2));
    orc_primes(2, N, pr, np);
    pr[np] = 0xFFFFFFFFu;
    size_t m = 0, i = 0;
    uint64_t acc = 0;
    int open = 0;
    for (uint32_t n = 2, k = 0; n <= N; ++n) {
        if (n == pr[k]) { ++k; continue; }
        if (!open) { acc = n; open = 1; continue; }
        u128 cc = (u128)acc * n;
        if (cc > 0xFFFFFFFFFFFFFFFFull) {
            if (m < cap) prod[m] = acc;
            ++m;
            acc = n;
        } else {
            acc = (uint64_t)cc;
        }
    }
    if (open) { if (m < cap) prod[m] = acc; ++m; }
    (void)i;
    free(pr);
    return m;
}

/* ---------------------------------------------------------------- one job */

typedef struct {
    uint32_t modes, nmin, nmax;
    /* factorial power table */
    size_t npow; uint64_t* pprod; uint32_t* pexp; uint32_t* pbit;
    /* primorial product table + iterated primes [nmin, nmax-1] */
    size_t nprod; uint64_t* prod; size_t niter; uint32_t* iter;
    /* compositorial product table + prime flags over [nmin, nmax) */
    size_t ncomp; uint64_t* comp; uint8_t* isprime_n;
} orc_tables;

static int build_tables(orc_tables* T, uint32_t modes, uint32_t nmin, uint32_t nmax) {
    memset(T, 0, sizeof(*T));
    T->modes = modes; T->nmin = nmin; T->nmax = nmax;
    uint32_t N = nmin - 1;
    if (modes & ORC_F) {
        T->npow = orc_power_table(N, NULL, NULL, NULL, 0);
        T->pprod = (uint64_t*)malloc(8 * (T->npow + 1));
        T->pexp = (uint32_t*)malloc(4 * (T->npow + 1));
        T->pbit = (uint32_t*)malloc(4 * (T->npow + 1));
        orc_power_table(N, T->pprod, T->pexp, T->pbit, T->npow);
    }
    if (modes & ORC_P) {
        T->nprod = orc_prime_products(N, NULL, 0);
        T->prod = (uint64_t*)malloc(8 * (T->nprod + 1));
        orc_prime_products(N, T->prod, T->nprod);
        T->niter = orc_primes(nmin, nmax - 1, NULL, 0);
        T->iter = (uint32_t*)malloc(4 * (T->niter + 1));
        orc_primes(nmin, nmax - 1, T->iter, T->niter);
        /* cl_sieve.cpp:1058-1066: pi(nmax) must equal pi(nmin-1) + |iter|,
         * i.e. nmax must not be prime. */
        size_t full = orc_primes(2, nmax, NULL, 0);
        size_t small = orc_primes(2, N, NULL, 0);
        if (full != small + T->niter) return -1;
    }
    if (modes & ORC_C) {
        T->ncomp = orc_comp_products(N, NULL, 0);
        T->comp = (uint64_t*)malloc(8 * (T->ncomp + 1));
        orc_comp_products(N, T->comp, T->ncomp);
        size_t w = (size_t)(nmax - nmin);
        T->isprime_n = (uint8_t*)calloc(w + 1, 1);
        size_t np = orc_primes(nmin, nmax, NULL, 0);
        uint32_t* pr = (uint32_t*)malloc(4 * (np + 1));
        orc_primes(nmin, nmax, pr, np);
        for (size_t i = 0; i < np; ++i)
            if (pr[i] < nmax) T->isprime_n[pr[i] - nmin] = 1;
        free(pr);
    }
    return 0;
}

static void free_tables(orc_tables* T) {
    free(T->pprod); free(T->pexp); free(T->pbit);
    free(T->prod); free(T->iter); free(T->comp); free(T->isprime_n);
}

/* Left-to-right powering of a Montgomery base by e, starting below the
 * leading bit (setup.cl:96-110). */
static uint64_t mont_pow_lr(uint64_t base, uint32_t e, uint32_t bit, const orc_mont* m) {
    if (e == 1) return base;
    uint64_t a = base;
    while (bit) {
        a = orc_mmul(a, a, m->p, m->q);
        if (e & bit) a = orc_mmul(a, base, m->p, m->q);
        bit >>= 1;
    }
    return a;
}

/* The 8-word record of one candidate after setup, walk and check: the
 * reference's ulong8 (getsegprimes.cl:221; iterate.cl:69,168). */
typedef struct { uint64_t s[8]; } orc_rec;

typedef struct { orc_factor* f; size_t n, cap; } fvec;
static void fpush(fvec* v, uint64_t p, int32_t nc, int32_t type) {
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 64;
        v->f = (orc_factor*)realloc(v->f, v->cap * sizeof(orc_factor));
    }
    v->f[v->n].p = p; v->f[v->n].nc = nc; v->f[v->n].type = type;
    v->n++;
}

/* Setup (setup.cl), walk (iterate.cl) and the checksum term (check.cl) for
 * one candidate p.  Returns the term; factor records go to fv. */
static uint64_t one_candidate(const orc_tables* T, uint64_t p, fvec* fv, orc_rec* rec) {
    orc_mont m = mont_init(p);
    const uint32_t modes = T->modes, nmin = T->nmin, nmax = T->nmax;
    uint64_t s4 = m.two, s6 = 0, s7 = 0;
    if (modes & (ORC_F | ORC_C)) s7 = orc_mmul(nmin - 1, m.r2, p, m.q);   /* setup.cl:75,182,233 */
    if (modes & ORC_F) {                                                   /* setup.cl:76-112 */
        uint32_t e = T->pexp[0], bit = T->pbit[0];
        uint64_t a = m.two;
        while (bit) {
            a = orc_mmul(a, a, p, m.q);
            if (e & bit) a = orc_madd(a, a, p);
            bit >>= 1;
        }
        s6 = a;
        for (size_t i = 1; i < T->npow; ++i) {
            uint64_t base = orc_mmul(T->pprod[i], m.r2, p, m.q);
            s6 = orc_mmul(s6, mont_pow_lr(base, T->pexp[i], T->pbit[i], &m), p, m.q);
        }
    }
    if (modes & ORC_P) {                                                   /* setup.cl:143-152 */
        for (size_t i = 0; i < T->nprod; ++i) {
            uint64_t mp = orc_mmul(T->prod[i], m.r2, p, m.q);
            s6 = i ? orc_mmul(s6, mp, p, m.q) : mp;
        }
    }
    uint64_t sc = 0;
    if (modes & ORC_C) {                                                   /* setup.cl:185-194, 247-282 */
        for (size_t i = 0; i < T->ncomp; ++i) {
            uint64_t mc = orc_mmul(T->comp[i], m.r2, p, m.q);
            sc = i ? orc_mmul(sc, mc, p, m.q) : mc;
        }
        if (!(modes & ORC_F)) s6 = sc; else s4 = sc;
    }

    /* walk */
    if (modes & ORC_P) {                                                   /* iterate.cl:102-111 */
        for (size_t j = 0; j < T->niter; ++j) {
            uint32_t qv = T->iter[j];
            uint64_t mq = orc_mmul(qv, m.r2, p, m.q);
            s6 = orc_mmul(s6, mq, p, m.q);
            if (s6 == m.one || s6 == m.nmo)
                fpush(fv, p, s6 == m.one ? -(int32_t)qv : (int32_t)qv, ORC_PRIMORIAL);
        }
    } else {
        const int F = (modes & ORC_F) != 0, C = (modes & ORC_C) != 0;
        uint64_t* sC = F ? &s4 : &s6;
        for (uint32_t n = nmin; n < nmax; ++n) {                            /* iterate.cl:72-80,135-147,173-191 */
            s7 = orc_madd(s7, m.one, p);
            if (F) {
                s6 = orc_mmul(s6, s7, p, m.q);
                if (s6 == m.one || s6 == m.nmo)
                    fpush(fv, p, s6 == m.one ? -(int32_t)n : (int32_t)n, ORC_FACTORIAL);
            }
            if (C && !T->isprime_n[n - nmin]) {
                *sC = orc_mmul(*sC, s7, p, m.q);
                if (*sC == m.one || *sC == m.nmo)
                    fpush(fv, p, *sC == m.one ? -(int32_t)n : (int32_t)n, ORC_COMPOSITORIAL);
            }
        }
    }
    if (rec) {
        rec->s[0] = p; rec->s[1] = m.q; rec->s[2] = m.r2; rec->s[3] = m.one;
        rec->s[4] = s4; rec->s[5] = m.nmo; rec->s[6] = s6; rec->s[7] = s7;
    }
    /* check.cl:59 (F or C: s6+s7), 115 (P: s6), 161 (F+C: s4+s6+s7) */
    if (modes & ORC_P) return s6;
    if ((modes & ORC_F) && (modes & ORC_C)) return s4 + s6 + s7;
    return s6 + s7;
}

/* One candidate's record, for unit checks of setup and walk. */
int orc_candidate_record(uint32_t modes, uint64_t p, uint32_t nmin, uint32_t nmax, uint64_t* rec8,
                         uint64_t* term) {
    orc_tables T;
    if (build_tables(&T, modes, nmin, nmax)) { free_tables(&T); return -1; }
    fvec fv = {0};
    orc_rec r;
    *term = one_candidate(&T, p, &fv, &r);
    memcpy(rec8, r.s, sizeof(r.s));
    free(fv.f);
    free_tables(&T);
    return 0;
}

static int factor_cmp(const void* a, const void* b) {
    const orc_factor* x = (const orc_factor*)a;
    const orc_factor* y = (const orc_factor*)b;
    if (x->p != y->p) return x->p < y->p ? -1 : 1;
    int32_t nx = x->nc < 0 ? -x->nc : x->nc, ny = y->nc < 0 ? -y->nc : y->nc;
    if (nx != ny) return nx < ny ? -1 : 1;
    return x->type < y->type ? -1 : (x->type > y->type);
}

/* Sort by (p, |n|) (cl_sieve.cpp:433-447; ties between an F and a C record of
 * the same (p, n), which the reference leaves to an unstable qsort, are put F
 * first), keep a factor when p passes isPrime (cl_sieve.cpp:561), and fold
 * n + c into the checksum (cl_sieve.cpp:583).  Kept factors are compacted to
 * the front; returns their number. */
size_t orc_harvest(orc_factor* f, size_t n, uint64_t* checksum) {
    qsort(f, n, sizeof(orc_factor), factor_cmp);
    uint64_t last = 0;
    size_t k = 0;
    for (size_t i = 0; i < n; ++i) {
        if (f[i].p == last || orc_is_prime(f[i].p)) {
            last = f[i].p;
            uint32_t nn = (uint32_t)(f[i].nc < 0 ? -f[i].nc : f[i].nc);
            int32_t c = f[i].nc < 0 ? -1 : 1;
            *checksum += (uint64_t)(uint32_t)(nn + (uint32_t)c);
            f[k++] = f[i];
        }
    }
    return k;
}

/* Whole job: candidates, setup, walk, check and harvest.  Returns 0, or -1
 * when the reference would stop with "CPU sieve failure" (primorial nmax
 * prime), -2 on an invalid mode set, -3 when the factor buffer is too small.
 * checksum/primecount/factorcount are the reference's final values
 * (cl_sieve.cpp:1748); dev_checksum is the device part alone (check.cl). */
int orc_run(uint32_t modes, uint64_t pmin, uint64_t pmax, uint32_t nmin, uint32_t nmax,
            orc_factor* fout, size_t fcap, size_t* nfac_raw, size_t* nfac, uint64_t* dev_checksum,
            uint64_t* checksum, uint64_t* primecount, uint64_t* factorcount) {
    if (!modes || ((modes & ORC_P) && (modes & (ORC_F | ORC_C)))) return -2;
    orc_tables T;
    if (build_tables(&T, modes, nmin, nmax)) { free_tables(&T); return -1; }
    size_t nc = orc_candidates(pmin, pmax, NULL, 0);
    uint64_t* cand = (uint64_t*)malloc(8 * (nc + 1));
    orc_candidates(pmin, pmax, cand, nc);
    uint64_t sum = 0;
    fvec all = {0};
#pragma omp parallel
    {
        fvec mine = {0};
        uint64_t s = 0;
#pragma omp for schedule(dynamic, 16)
        for (int64_t i = 0; i < (int64_t)nc; ++i) s += one_candidate(&T, cand[i], &mine, NULL);
#pragma omp critical
        {
            sum += s;
            for (size_t k = 0; k < mine.n; ++k) fpush(&all, mine.f[k].p, mine.f[k].nc, mine.f[k].type);
        }
        free(mine.f);
    }
    free(cand);
    free_tables(&T);
    *dev_checksum = sum;
    *primecount = nc;
    *nfac_raw = all.n;
    uint64_t cs = sum;
    size_t kept = orc_harvest(all.f, all.n, &cs);
    *checksum = cs;
    *factorcount = kept;
    *nfac = kept;
    int rc = 0;
    if (kept > fcap) rc = -3;
    else if (kept) memcpy(fout, all.f, kept * sizeof(orc_factor));
    free(all.f);
    return rc;
}

/* ---------------------------------------------------------------- verify */

/* n!, n# or n!/n# mod p by direct products (the slow check of
 * verifyprime.cpp:149-218, restated from 1 rather than from the 34!, 101#,
 * 44!/44# constants) and the +-1 test. */
int orc_verify(uint64_t p, uint32_t n, int32_t c, int32_t type) {
    uint64_t r = 1 % p;
    if (type == ORC_FACTORIAL) {
        for (uint32_t i = 2; i <= n && r; ++i) r = (uint64_t)(((u128)r * i) % p);
    } else {
        size_t np = orc_primes(2, n, NULL, 0);
        uint32_t* pr = (uint32_t*)malloc(4 * (np + 2));
        orc_primes(2, n, pr, np);
        pr[np] = 0xFFFFFFFFu;
        if (type == ORC_PRIMORIAL) {
            for (size_t i = 0; i < np && r; ++i) r = (uint64_t)(((u128)r * pr[i]) % p);
        } else {
            for (uint32_t i = 2, k = 0; i <= n && r; ++i) {
                if (i == pr[k]) { ++k; continue; }
                r = (uint64_t)(((u128)r * i) % p);
            }
        }
        free(pr);
    }
    if (r == 1 && c == -1) return 1;
    if (r == p - 1 && c == 1) return 1;
    return 0;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
