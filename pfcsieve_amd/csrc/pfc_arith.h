// pfc_arith.h -- exact modular arithmetic for 64-bit odd moduli p, shared by
// the gfx950 kernels and the host (the same source compiles for both, so the
// host build can check the device arithmetic against big integers).
//
// Representation ("shifted"): for a prime p, s = clz64(p) and d = p << s, so
// d has its top bit set.  A residue x mod p is kept as x' = x << s, i.e. as a
// residue mod d that is a multiple of 2^s.  Multiplying x' by a plain integer
// n and reducing mod d gives (x*n mod p) << s, so the shift is preserved.
//
// Reduction is the 3-by-2 division with a precomputed reciprocal
//   v = floor((2^96 - 1) / d) - 2^32
// (N. Moller and T. Granlund, "Improved division by invariant integers",
// IEEE Trans. Computers 60(2), 2011, Algorithm 5), with the remainder formed
// from per-modulus constants (rem3by2_nofix).  A 64x32-bit modular product
// costs five 32x32+64 multiply-adds (v_mad_u64_u32 on CDNA).
//
// The reference (kernels/iterate.cl:44-57) multiplies 64-bit Montgomery forms
// with R = 2^64 and its checksum sums Montgomery forms X*R mod p.  Here the
// walk keeps plain residues X << s, so that +1 and -1 are 2^s and d - 2^s,
// the two ends of the residue range, and the checksum term X*R mod p is
// formed once per candidate at the end.  Both give the same canonical
// integers (residues, factor records and checksum terms).
#ifndef PFC_ARITH_H
#define PFC_ARITH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define PFC_HD __host__ __device__ __forceinline__
#else
#define PFC_HD static inline
#endif

namespace pfc {

struct Modulus {
    uint32_t d1, d0;  // d = p << s, high and low words (d1 >= 2^31)
    uint32_t v;       // 3-by-2 reciprocal of d
    uint32_t s;       // shift, clz64(p)
};

PFC_HD uint64_t mod_d(const Modulus& m) { return ((uint64_t)m.d1 << 32) | m.d0; }

PFC_HD uint32_t clz64(uint64_t x) { return (uint32_t)__builtin_clzll(x); }

// 128-bit product high word of two 64-bit values, from 32-bit pieces.
PFC_HD uint64_t mulhi64(uint64_t a, uint64_t b) {
    uint64_t a0 = (uint32_t)a, a1 = a >> 32, b0 = (uint32_t)b, b1 = b >> 32;
    uint64_t p00 = a0 * b0, p01 = a0 * b1, p10 = a1 * b0, p11 = a1 * b1;
    uint64_t mid = (p00 >> 32) + (uint32_t)p01 + (uint32_t)p10;
    return p11 + (p01 >> 32) + (p10 >> 32) + (mid >> 32);
}

// Remainder of <u2,u1,u0> by d, given hi = <u2,u1> < d, without the final
// (rare) correction: returns x with x == r or x == r + d for the remainder r,
// and sets *fix when the caller must subtract d (x >= d).  The kernels use
// this form so that the rare correction becomes a wavefront-uniform branch.
//
// Algorithm 5 of Moller-Granlund: <q1,q0> = v*u2 + <u2,u1> estimates the
// quotient; r = <u1,u0> - (q1 + 1)*d (mod 2^64) is the remainder for quotient
// q1 + 1, d is added back when the high word of r is >= q0 (the quotient was
// one too large), and a last, unlikely, subtraction of d covers a quotient
// one too small.  This step forms r = <u1,u0> - (q1 + 1)*d with per-modulus
// constants instead of a borrow chain: with k1 = -(d1 + 1) and k0 = 2^32 - d0
// (k1 = -d1 and k0 = 0 when d0 == 0) and nd = 2^64 - d,
//   r = q1 * k0 + (<u1 + q1 * k1, u0> + nd)      (mod 2^64),
// since q1 * k0 = q1 * 2^32 - q1 * d0 and the extra q1 * 2^32 cancels the -q1
// folded into k1.  u1 + q1 * k1 is the low word of one multiply-add whose
// 64-bit addend is <u2,u1> as computed, and the two remaining terms are one
// 64-bit add and one multiply-add.
PFC_HD uint64_t rem3by2_nofix(uint64_t hi, uint32_t u0, uint32_t k1, uint32_t k0, uint64_t nd, uint64_t d,
                              uint32_t v, bool* fix) {
    const uint32_t u2 = (uint32_t)(hi >> 32);
    const uint64_t qq = (uint64_t)v * u2 + hi;
    const uint32_t q1 = (uint32_t)(qq >> 32), q0 = (uint32_t)qq;
    const uint32_t w = (uint32_t)((uint64_t)q1 * k1 + hi);              // u1 - q1 * (d1 + 1)
    const uint64_t pair = (((uint64_t)w) << 32) | u0;                   // <w, u0>
#if defined(__HIP_DEVICE_COMPILE__)
    // One 64-bit add on the assembled pair.  Written as an instruction so the
    // optimiser cannot split the pair and reassociate the sum into two adds.
    uint64_t y;
    asm("v_lshl_add_u64 %0, %1, 0, %2" : "=v"(y) : "v"(pair), "v"(nd));
#else
    const uint64_t y = pair + nd;
#endif
    uint64_t x = (uint64_t)q1 * k0 + y;
#if defined(__HIP_DEVICE_COMPILE__)
    // An exec-masked add instead of two selects: the empty volatile asm keeps
    // the compiler from turning the branch into v_cndmask pairs, so the
    // correction costs one VALU add plus scalar exec-mask bookkeeping.
    if ((uint32_t)(x >> 32) >= q0) {
        asm volatile("" ::: );
        x += d;
    }
#else
    if ((uint32_t)(x >> 32) >= q0) x += d;
#endif
    *fix = x >= d;
    return x;
}

// (a * n) mod d for a < d and any 32-bit n: the 96-bit product <u2,u1,u0>,
// then rem3by2_nofix (the same contract: x or x + d, *fix when x >= d).
PFC_HD uint64_t mulmod32_nofix(uint64_t a, uint32_t n, uint32_t k1, uint32_t k0, uint64_t nd, uint64_t d,
                               uint32_t v, bool* fix) {
    const uint64_t lo = (uint64_t)(uint32_t)a * n;
    const uint64_t hi = (uint64_t)(uint32_t)(a >> 32) * n + (lo >> 32);   // <u2,u1>
    return rem3by2_nofix(hi, (uint32_t)lo, k1, k0, nd, d, v, fix);
}

// The walk's test for a result x of mulmod32_nofix against a band B (B = 1
// shifted, 2^s, or B = b << s for a group of steps): it must pass every x
// with x <= B or x >= d - B (a hit, or the residue after a hit at an untested
// step) and every unreduced x (x >= d).  It looks at the high word alone:
// with bh = B >> 32 and dh = (d - B) >> 32, x <= B implies x_hi <= bh and
// x >= d - B implies x_hi >= dh, so
//   x_hi + c >= thr  (mod 2^32),  c = ~bh = -bh - 1,  thr = dh - bh - 1,
// (x_hi - bh - 1 wraps for x_hi <= bh, and is at least thr exactly when
// x_hi >= dh) passes all of them with one 32-bit add and one 32-bit compare.
// It also passes the x with x_hi == bh above B or x_hi == dh below d - B, a
// fraction below 2^33 / d of the residues, which the slow path re-tests
// exactly.  When dh <= bh the band covers the high words of [B, d - B)
// entirely and the constants make every x pass (c = thr = 0).
PFC_HD void walk_band(uint64_t d, uint64_t B, uint32_t* c, uint32_t* thr) {
    const uint32_t bh = (uint32_t)(B >> 32), dh = (uint32_t)((d - B) >> 32);
    if (dh > bh) {
        *c = ~bh;
        *thr = dh - bh - 1;
    } else {
        *c = 0;
        *thr = 0;
    }
}
PFC_HD bool walk_attention(uint64_t x, uint32_t c, uint32_t thr) { return (uint32_t)(x >> 32) + c >= thr; }

// The constants k1, k0 of mulmod32_nofix for d = <d1, d0>.
PFC_HD uint32_t walk_k1(uint32_t d1, uint32_t d0) { return d0 ? ~d1 : 0u - d1; }
PFC_HD uint32_t walk_k0(uint32_t d0) { return 0u - d0; }

// ---------------------------------------------------------------- narrow walk

// Walk arithmetic for moduli with 2^34 < p < 2^55 and multipliers n with
// p * n < 2^62 and n < 2^22 ("narrow" moduli; the headline's p ~ 10^12 with n
// below 10^6 is one).  Residues are plain integers X in [0, 2p), not reduced
// and not shifted.  Then t = X * n < 2^63 is one 64-bit word and a quotient
// estimate from its high word alone leaves a remainder below 2p, so a step is
// five multiplies and no correction:
//   t  = X0 * n + ((X1 * n) << 32)     X1 * n < 2^31 (one multiply, one
//                                      multiply-add on the pair <X1*n, 0>)
//   q  = hi32(t1 * m)                  m = floor(2^64 / p)
//   r  = t - q * p = q * np0 + <t1 - q * (p1 + 1), t0>   (mod 2^64)
// with t1, t0 the words of t, p1 the high word of p and np0 = 2^32 - p0
// (p0 != 0: p is odd).  The correction term is exact: with
// 2^64 / p = m + phi, 0 <= phi < 1 (m is the floor, from the exact 3-by-2
// reciprocal),
//   t / p - q < 1 + t1 * phi / 2^32 + t0 / p < 1 + 2^-1 + 2^-2,
// because t1 < 2^31 and t0 < 2^32 < p / 4; so floor(t / p) <= q + 1 and
// 0 <= r < 2p (q <= t / p since m <= 2^64 / p).  q < 2n < 2^23 and p1 + 1 <= 2^23, so q * (p1 + 1) fits the
// signed 24-bit multiply-add (v_mad_i32_i24) on the device.
PFC_HD bool narrow_fits(uint64_t p, uint32_t bmax) {
    if (bmax == 0 || bmax >= (1u << 22) || p <= (1ull << 34) || p >= (1ull << 55)) return false;
    const uint64_t a = (p >> 32) * bmax;                       // < 2^45
    if (a >= (1ull << 30)) return false;
    return (a << 32) + (uint64_t)(uint32_t)p * bmax < (1ull << 62);
}

// floor(2^64 / p) from the 3-by-2 reciprocal of d = p << s (s < 30 here):
// v + 2^32 = floor((2^96 - 1) / d) = floor(2^96 / d) (d does not divide 2^96),
// and floor(floor(x) / 2^(32-s)) = floor(x / 2^(32-s)) = floor(2^64 / p).
PFC_HD uint32_t narrow_m(uint32_t v, uint32_t s) { return (uint32_t)(((1ull << 32) + v) >> (32 - s)); }

// The per-modulus constants of narrow_mul.
PFC_HD uint32_t narrow_np0(uint64_t p) { return 0u - (uint32_t)p; }
PFC_HD uint32_t narrow_kp1(uint64_t p) { return 0u - ((uint32_t)(p >> 32) + 1u); }

PFC_HD uint32_t mulhi32(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umulhi(a, b);
#else
    return (uint32_t)(((uint64_t)a * b) >> 32);
#endif
}

// X * n mod p up to one p, for X < 2p: a value in [0, 2p) congruent to X * n.
PFC_HD uint64_t narrow_mul(uint64_t x, uint32_t n, uint32_t m, uint32_t np0, uint32_t kp1) {
    uint32_t a = (uint32_t)(x >> 32) * n;
#if defined(__HIP_DEVICE_COMPILE__)
    // The empty asm statements keep the optimiser from recognising the
    // 64-by-32 product and rebuilding it as a 96-bit one (with register
    // moves around it): t is the multiply-add on the pair <a, 0>, and w lands
    // in t's high register, so that <w, t0> is again a pair.
    asm volatile("" : "+v"(a));
#endif
    uint64_t t = (uint64_t)(uint32_t)x * n + ((uint64_t)a << 32);
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+v"(t));
#endif
    const uint32_t t1 = (uint32_t)(t >> 32);
    const uint32_t q = mulhi32(t1, m);
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t w = (uint32_t)(__mul24((int)q, (int)kp1) + (int)t1);
    asm volatile("" : "+v"(w));
#else
    const uint32_t w = t1 + q * kp1;
#endif
    return (uint64_t)q * np0 + (((uint64_t)w << 32) | (uint32_t)t);
}

// The narrow walk's test of a reduced residue x in [0, p) against a band B:
// pass iff x <= B or x >= p - B, as x + c >= thr with c = 2^64 - B - 1 and
// thr = p - 2B - 1 (one 64-bit add and one 64-bit compare; needs p > 2B + 1).
PFC_HD void narrow_band(uint64_t p, uint64_t B, uint64_t* c, uint64_t* thr) {
    *c = ~B;
    *thr = p - 2 * B - 1;
}
PFC_HD bool narrow_attention(uint64_t x, uint64_t c, uint64_t thr) { return x + c >= thr; }

// ---------------------------------------------------------------- medium walk

// Walk arithmetic for 2^32 < p < 2^63 and multipliers n < 2^28 ("medium"
// moduli, the ones outside the narrow form but for the top bit).  Residues
// are kept at half the shifted scale: D = p << (s - 1) = d / 2 lies in
// (2^62, 2^63), and x in [0, 2D) stands for X = x >> (s - 1).  With
// M = floor(2^94 / D) (in (2^31, 2^32)), the quotient estimate
//   f = hi32(x1 * M),   q = hi32(f * 4n)        (x1 = x >> 32)
// is floor(x * n / D) or one less: f <= x * 2^30 / D < f + 3 (x's low word,
// M's truncation and the floor each lose less than one), so
//   x * n / D - q < 1 + 3n / 2^30 < 2   for n < 2^28.
// The remainder r = x * n - q * D < 2D needs only the low 64 bits of x * n:
//   T = x0 * n + ((x1 * n) << 32),  r = q * nD0 + <T1 - q * (D1 + 1), T0>
// (mod 2^64), with nD0 = 2^32 - D0 (D0 != 0: its lowest set bit is bit
// s - 1 < 31).  Seven instructions per step: v_mul_lo_u32 and v_mad_u64_u32
// (T), two v_mul_hi_u32 (f, q), v_mul_lo_u32 and v_add_u32 (T1 - q(D1+1)),
// v_mad_u64_u32 (r).
PFC_HD bool medium_fits(uint64_t p, uint32_t bmax) {
    return bmax != 0 && bmax < (1u << 28) && p > (1ull << 32) && p < (1ull << 63);
}

// M = floor(2^94 / D) from the 3-by-2 reciprocal of d = 2D:
// v + 2^32 = floor(2^96 / d) = floor(2^95 / D), and halving it floors again.
PFC_HD uint32_t medium_m(uint32_t v) { return (uint32_t)(((1ull << 32) + v) >> 1); }

// x * n mod D up to one D, for x < 2D and n < 2^28 (n4 = 4n): a value in
// [0, 2D) congruent to x * n; k1 = -(D1 + 1) mod 2^32, nd0 = 2^32 - D0.
PFC_HD uint64_t medium_mul(uint64_t x, uint32_t n, uint32_t n4, uint32_t M, uint32_t nd0, uint32_t k1) {
    const uint32_t x1 = (uint32_t)(x >> 32);
    uint32_t a = x1 * n;
#if defined(__HIP_DEVICE_COMPILE__)
    // As in narrow_mul: T is the multiply-add on the pair <a, 0>, and the
    // corrected high word is written back into T's high register.
    asm volatile("" : "+v"(a));
#endif
    uint64_t T = (uint64_t)(uint32_t)x * n + ((uint64_t)a << 32);    // x * n mod 2^64
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+v"(T));
#endif
    const uint32_t f = mulhi32(x1, M);
    const uint32_t q = mulhi32(f, n4);
#if defined(__HIP_DEVICE_COMPILE__)
    // T + ((q * k1) << 32) as one 64-bit add on the pair <q * k1, 0> (built
    // like <a, 0> above); left to itself the optimiser fuses q * k1 into a
    // 64-bit multiply-add and reassembles the words with register moves.
    uint32_t qk = q * k1;
    asm volatile("" : "+v"(qk));
    uint64_t Tw;
    asm("v_lshl_add_u64 %0, %1, 0, %2" : "=v"(Tw) : "v"(T), "v"((uint64_t)qk << 32));
    return (uint64_t)q * nd0 + Tw;
#else
    const uint32_t w = (uint32_t)(T >> 32) + q * k1;
    return (uint64_t)q * nd0 + (((uint64_t)w << 32) | (uint32_t)T);
#endif
}

// (a * n) mod d for a < d and any 32-bit n.  Keeps the shifted form.
PFC_HD uint64_t mulmod32(uint64_t a, uint32_t n, const Modulus& m) {
    bool fix;
    const uint64_t d = mod_d(m);
    const uint64_t x = mulmod32_nofix(a, n, walk_k1(m.d1, m.d0), walk_k0(m.d0), 0ull - d, d, m.v, &fix);
    return fix ? x - d : x;
}

// (a + b) mod d for a, b < d (d >= 2^63, so a + b may carry out of 64 bits).
PFC_HD uint64_t addmod(uint64_t a, uint64_t b, const Modulus& m) {
    const uint64_t d = mod_d(m);
    const uint64_t r = a + b;
    return (r < a || r >= d) ? r - d : r;
}

// (a * b) mod d up to one d, for plain integers a, b with a * b < d * 2^64
// (a < d, or b < 2^63 and any a): the 128-bit product <t3,t2,t1,t0>, then
// two 3-by-2 steps, <t3,t2,t1> mod d and <r,t0> mod d (the first step's
// remainder is corrected so that it satisfies the second's <u2,u1> < d
// precondition).  Returns x == r or r + d for the remainder r and sets *fix
// when x >= d, like rem3by2_nofix.  On the device the rare correction of the
// first step is a wavefront-uniform branch.
PFC_HD uint64_t mulmod64_nofix(uint64_t a, uint64_t b, const Modulus& m, bool* fix) {
    const uint64_t d = mod_d(m), nd = 0ull - d;
    const uint32_t k1 = walk_k1(m.d1, m.d0), k0 = walk_k0(m.d0);
    const uint64_t a0 = (uint32_t)a, a1 = a >> 32, b0 = (uint32_t)b, b1 = b >> 32;
    const uint64_t p00 = a0 * b0;
    const uint64_t p01 = a0 * b1 + (p00 >> 32);
    const uint64_t p10 = a1 * b0 + (uint32_t)p01;
    const uint64_t p11 = a1 * b1 + (p01 >> 32) + (p10 >> 32);
    bool f1;
    uint64_t r = rem3by2_nofix(p11, (uint32_t)p10, k1, k0, nd, d, m.v, &f1);
#if defined(__HIP_DEVICE_COMPILE__)
    if (__builtin_expect(__ballot(f1) != 0, 0))
#endif
        if (r >= d) r -= d;
    return rem3by2_nofix(r, (uint32_t)p00, k1, k0, nd, d, m.v, fix);
}

// (a * b) mod d for a < d and any 64-bit b (b is a plain integer).
PFC_HD uint64_t mulmod64(uint64_t a, uint64_t b, const Modulus& m) {
    bool f;
    uint64_t r = mulmod64_nofix(a, b, m, &f);
#if defined(__HIP_DEVICE_COMPILE__)
    if (__builtin_expect(__ballot(f) != 0, 0))
#endif
        if (f) r -= mod_d(m);
    return r;
}

// Square of the residue a' = x << s: (x*x mod p) << s.
PFC_HD uint64_t sqrmod(uint64_t a, const Modulus& m) { return mulmod64(a, a >> m.s, m); }

// Builds the modulus for odd p >= 3.  The reciprocal comes from a double
// estimate corrected against exact 128-bit products d*V <= 2^96-1 < d*(V+1).
PFC_HD Modulus make_modulus(uint64_t p) {
    Modulus m;
    m.s = clz64(p);
    const uint64_t d = p << m.s;
    m.d1 = (uint32_t)(d >> 32);
    m.d0 = (uint32_t)d;
    // 2^96 / d lies in (2^32, 2^33); the double estimate is within a few units.
    uint64_t V = (uint64_t)(79228162514264337593543950336.0 / (double)d);
    // d*V as a 128-bit number compared with 2^96 - 1 = <0xFFFFFFFF, 2^64-1>.
    for (;;) {
        const uint64_t hi = mulhi64(d, V);
        if (hi > 0xFFFFFFFFull) { --V; continue; }
        const uint64_t lo = d * (V + 1);
        const uint64_t hi1 = mulhi64(d, V + 1);
        (void)lo;
        if (hi1 <= 0xFFFFFFFFull) { ++V; continue; }
        break;
    }
    m.v = (uint32_t)(V - (1ull << 32));
    return m;
}

// R mod p in shifted form, R = 2^64: 2^(64+s) mod d.  2^64 mod d = 2^64 - d
// because 2^63 < d < 2^64; the remaining factor 2^s is applied in two steps
// of at most 2^31 each (s <= 62 for p >= 3).
PFC_HD uint64_t mont_one(const Modulus& m) {
    const uint64_t x = 0 - mod_d(m);
    const uint32_t s1 = m.s < 31 ? m.s : 31;
    return mulmod32(mulmod32(x, 1u << s1, m), 1u << (m.s - s1), m);
}

// True when d*(v + 2^32) <= 2^96 - 1 < d*(v + 2^32 + 1), i.e. the reciprocal
// is exact for d.  Used as the device-side state check.
PFC_HD bool modulus_ok(uint64_t p, const Modulus& m) {
    if (!(p & 1) || m.s != clz64(p) || mod_d(m) != (p << m.s)) return false;
    const uint64_t d = mod_d(m);
    const uint64_t V = (uint64_t)m.v + (1ull << 32);
    return mulhi64(d, V) <= 0xFFFFFFFFull && mulhi64(d, V + 1) > 0xFFFFFFFFull;
}

// Base-2 strong probable-prime test of odd N >= 3 (the test of
// kernels/getsegprimes.cl:79-108, computed in the shifted representation).
PFC_HD bool strong_prp2(uint64_t N) {
    const Modulus m = make_modulus(N);
    const uint64_t d = mod_d(m);
    const uint64_t one = 1ull << m.s;          // 1 in shifted form
    const uint64_t nmo = d - one;              // N - 1 in shifted form
    const uint32_t t = (uint32_t)__builtin_ctzll(N - 1);
    const uint64_t e = N >> t;                 // odd part of N - 1
    // Bit below the leading one of e.  For e == 1 (N = 2^t + 1) this is 0 and
    // the test starts from 2^1; the reference's 6-bit-masked shift instead
    // raises 2 to 2^64+1, which gives the same verdict for every such N below
    // 2^64 that survives the 7..113 sieve (257, 65537, 2^32+1, 2^40+1,
    // 2^48+1, 2^56+1; checked in the tests).
    uint64_t bit = (1ull << (63 - clz64(e))) >> 1;
    uint64_t a = 2ull << m.s;                  // 2 (N >= 3, so 2 < N)
    while (bit) {
        a = sqrmod(a, m);
        if (e & bit) a = addmod(a, a, m);
        bit >>= 1;
    }
    if (a == one || a == nmo) return true;
    for (uint32_t i = 1; i < t; ++i) {
        a = sqrmod(a, m);
        if (a == nmo) return true;
    }
    return false;
}

// Base-2 strong probable-prime test of K odd values N >= 3 in lockstep, with
// no data-dependent branch: for M = N - 1 and t = ctz(M), a = 2^(M >> i) mod N
// is built over the bits of M from the top (a <- a^2, then a <- 2a where bit
// i of M is set; while i is above M's leading bit, a stays 1, since 1^2 = 1),
// so the trip count is the same for every value of bit length up to `bits`.
// At i = t, a = 2^e with e the odd part of M (the strong test's x0, which
// passes as +-1); at each i from t - 1 down to 1, a = x0^(2^(t-i)), which
// passes as -1.  `tmax` >= every t: the rounds above it skip the checks.  For
// N = 2^t + 1 (e = 1) the test starts from 2^1, as strong_prp2 does.  On the
// device `bits` and `tmax` are wavefront-uniform, so K candidates per lane run
// as K interleaved squaring chains.
template <int K>
PFC_HD void strong_prp2_lockstep(const uint64_t (&N)[K], uint32_t bits, uint32_t tmax, bool (&prp)[K]) {
    Modulus m[K];
    uint64_t d[K], one[K], nmo[K], M[K], a[K];
    uint32_t t[K];
    for (int k = 0; k < K; ++k) {
        m[k] = make_modulus(N[k]);
        d[k] = mod_d(m[k]);
        one[k] = 1ull << m[k].s;
        nmo[k] = d[k] - one[k];
        M[k] = N[k] - 1;
        t[k] = (uint32_t)__builtin_ctzll(M[k]);
        a[k] = one[k];
        prp[k] = false;
    }
    int i = (int)bits - 1;
    for (; i > (int)tmax; --i) {
        for (int k = 0; k < K; ++k) {
            a[k] = mulmod64(a[k], a[k] >> m[k].s, m[k]);
            const uint64_t a2 = addmod(a[k], a[k], m[k]);
            a[k] = ((M[k] >> i) & 1) ? a2 : a[k];
        }
    }
    for (; i >= 1; --i) {
        for (int k = 0; k < K; ++k) {
            a[k] = mulmod64(a[k], a[k] >> m[k].s, m[k]);
            const uint64_t a2 = addmod(a[k], a[k], m[k]);
            a[k] = ((M[k] >> i) & 1) ? a2 : a[k];
            const uint32_t ui = (uint32_t)i;
            prp[k] |= (ui <= t[k] && a[k] == nmo[k]) || (ui == t[k] && a[k] == one[k]);
        }
    }
}

// ---------------------------------------------------------------- 32-bit primality

// Montgomery product a*b*2^-32 mod n for odd n < 2^32 and a, b < n;
// nneg = -n^-1 mod 2^32.
PFC_HD uint32_t mont32_mul(uint32_t a, uint32_t b, uint32_t n, uint32_t nneg) {
    const uint64_t t = (uint64_t)a * b;
    const uint32_t q = (uint32_t)t * nneg;
    const uint64_t qn = (uint64_t)q * n;
    // t + qn == 0 mod 2^32; the low words add to 0 or 2^32.
    const uint64_t carry = ((uint64_t)(uint32_t)t + (uint32_t)qn) >> 32;
    const uint64_t r = (t >> 32) + (qn >> 32) + carry;   // < 2n
    return (uint32_t)(r >= n ? r - n : r);
}

// Deterministic primality of a 32-bit integer: strong probable-prime tests to
// the bases 2, 7 and 61 (Jaeschke 1993: the least composite passing all
// three is 4,759,123,141 > 2^32).  Used by the start-table self-check as a
// primality source independent of the host sieve.
PFC_HD bool is_prime_u32(uint32_t n) {
    if (n < 2) return false;
    if (n < 4) return true;
    if (!(n & 1u)) return false;
    uint32_t inv = n;                          // n*n == 1 mod 8: 3 correct bits
    for (int i = 0; i < 4; ++i) inv *= 2u - n * inv;
    const uint32_t nneg = 0u - inv;
    const uint32_t one = (uint32_t)(0x100000000ull % n);   // 2^32 mod n
    const uint32_t mone = n - one;                          // -1
    uint32_t d = n - 1;
    uint32_t t = 0;
    while (!(d & 1u)) {
        d >>= 1;
        ++t;
    }
    const uint32_t bases[3] = {2u, 7u, 61u};
    for (int b = 0; b < 3; ++b) {
        const uint32_t a = bases[b] % n;
        if (a == 0) continue;
        const uint32_t am = (uint32_t)(((uint64_t)a << 32) % n);
        uint32_t x = one;
        for (int bit = 31 - __builtin_clz(d); bit >= 0; --bit) {
            x = mont32_mul(x, x, n, nneg);
            if ((d >> bit) & 1u) x = mont32_mul(x, am, n, nneg);
        }
        if (x == one || x == mone) continue;
        bool witness = true;
        for (uint32_t i = 1; i < t && witness; ++i) {
            x = mont32_mul(x, x, n, nneg);
            if (x == mone) witness = false;
        }
        if (witness) return false;
    }
    return true;
}

// ---------------------------------------------------------------- start residues

// Start residue (X mod p) << s of X = prod over primes r <= N of r^e(r), from
// a table that writes the exponents in base B = 2^k (k = digit_bits, 1..8):
// group (b, t) holds the chunks (products of primes below 2^63) of the primes
// whose exponent digit at position b is t, and
//   X = prod_b H_b^(B^b),   H_b = prod_{t=1..B-1} D_{b,t}^t,
// with D_{b,t} the product of group (b, t).  Horner's rule over the positions,
// high first, takes acc <- acc^B * H_b, and H_b costs each chunk once plus
// one product per digit by the suffix products S_t = prod_{t' >= t} D_{b,t'}:
// H_b = prod_t S_t.  Groups are stored position-major from the highest
// position, digits from B-1 down to 1: group g = (npos-1-b)*(B-1) + (B-1-t),
// chunks groups[g] .. groups[g+1].  For k = 1 this is the plain binary
// Horner scheme (acc <- acc^2 * D_b), evaluated without the suffix product.
//
// Chunk products (mulmod64_nofix) accept the unreduced accumulator, which is
// reduced only before a squaring or a residue-by-residue product and at the
// end.  With first/stride, only the chunks first, first + stride, ... of each
// group are taken: X is multiplicative in the chunks for every exponent
// layout, so the product over first = 0 .. stride - 1 of those partial values
// is the full one, which lets a workgroup share the work.
template <class Chunks, class Groups>
PFC_HD uint64_t start_residue(Chunks chunks, Groups groups, uint32_t npos, uint32_t k, const Modulus& m,
                              uint32_t first = 0, uint32_t stride = 1) {
    const uint64_t d = mod_d(m);
    const uint64_t one = 1ull << m.s;                               // 1, shifted
    const uint32_t ndig = (1u << k) - 1;
    uint64_t acc = one;
    uint32_t g = 0;
    bool fix;
    for (int b = (int)npos - 1; b >= 0; --b) {
        if (b != (int)npos - 1) {
            for (uint32_t i = 0; i < k; ++i) {
                if (acc >= d) acc -= d;
                acc = mulmod64(acc, acc >> m.s, m);
            }
        }
        if (ndig == 1) {
            const uint32_t g0 = groups[g], g1 = groups[g + 1];
            for (uint32_t i = g0 + first; i < g1; i += stride) acc = mulmod64_nofix(acc, chunks[i], m, &fix);
            ++g;
            continue;
        }
        uint64_t S = one;
        bool started = false;
        for (uint32_t t = 0; t < ndig; ++t, ++g) {
            const uint32_t g0 = groups[g], g1 = groups[g + 1];
            for (uint32_t i = g0 + first; i < g1; i += stride) S = mulmod64_nofix(S, chunks[i], m, &fix);
            started |= g1 > g0;
            if (started) {
                // both canonical: acc * (S >> s) < d * p <= d * 2^64
                if (S >= d) S -= d;
                if (acc >= d) acc -= d;
                acc = mulmod64_nofix(acc, S >> m.s, m, &fix);
            }
        }
    }
    return acc >= d ? acc - d : acc;
}

// Montgomery form (x * 2^64 mod p) << s of a canonical shifted residue x,
// given mone = mont_one(m): the checksum terms of the reference.
PFC_HD uint64_t to_mont(uint64_t x, uint64_t mone, const Modulus& m) {
    return mulmod64(x, mone >> m.s, m);
}

// ---------------------------------------------------------------- wheel + small-prime sieve

// The sieve primes of the reference's candidate generator (kernels/
// getsegprimes.cl:116-168, 194-197): every prime from 7 to 113.
constexpr uint32_t kSievePrimes[27] = {7,  11, 13, 17, 19, 23, 29, 31,  37,  41,  43,  47,  53, 59,
                                       61, 67, 71, 73, 79, 83, 89, 97, 101, 103, 107, 109, 113};

// Bit j set iff q divides j, for 0 <= j < 64.
PFC_HD constexpr uint64_t sieve_pattern(uint32_t q) {
    uint64_t m = 0;
    for (uint32_t j = 0; j < 64; j += q) m |= 1ull << j;
    return m;
}

// Offsets 0 <= j < 60 with start + j coprime to 30, for start a multiple of 30.
constexpr uint64_t kWheel60 = (1ull << 1) | (1ull << 7) | (1ull << 11) | (1ull << 13) |
                              (1ull << 17) | (1ull << 19) | (1ull << 23) | (1ull << 29) |
                              (1ull << 31) | (1ull << 37) | (1ull << 41) | (1ull << 43) |
                              (1ull << 47) | (1ull << 49) | (1ull << 53) | (1ull << 59);

// Survivors of the 60-wide window [start, start + 60), start a multiple of
// 30: bit j is set iff lo <= start + j < hi, start + j is coprime to 30 and
// has no prime factor from 7 to 113.  rmod[i] = start mod kSievePrimes[i].
PFC_HD uint64_t window_survivors(uint64_t start, uint64_t lo, uint64_t hi, const uint32_t* rmod) {
    uint64_t cand = kWheel60;
    if (start < lo) {
        const uint64_t skip = lo - start;
        cand = skip >= 64 ? 0 : cand & (~0ull << skip);
    }
    if (hi <= start) {
        cand = 0;
    } else if (hi - start < 64) {
        cand &= (1ull << (hi - start)) - 1;
    }
    if (!cand) return 0;
    uint64_t hit = 0;
    for (int i = 0; i < 27; ++i) {
        const uint32_t q = kSievePrimes[i];
        const uint32_t j0 = rmod[i] ? q - rmod[i] : 0;   // offset of the first multiple of q
        // A shift count of 64 or more is out of range (the hardware would use
        // it mod 64), and such a multiple lies outside the window anyway.
        if (j0 < 64) hit |= sieve_pattern(q) << j0;
    }
    return cand & ~hit;
}

}  // namespace pfc

#endif  // PFC_ARITH_H
