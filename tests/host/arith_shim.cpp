// Host build of the device arithmetic (pfcsieve_amd/csrc/pfc_arith.h) so the
// CPU tests can check it against Python big integers.  Test-only.
#include <stddef.h>
#include <stdint.h>
#include "../../pfcsieve_amd/csrc/pfc_arith.h"
#include "../../pfcsieve_amd/csrc/pfc_share.h"

extern "C" {
void t_modulus(uint64_t p, uint32_t out[4]) {
    pfc::Modulus m = pfc::make_modulus(p);
    out[0] = m.d1; out[1] = m.d0; out[2] = m.v; out[3] = m.s;
}
uint64_t t_mulmod32(uint64_t p, uint64_t a_shifted, uint32_t n) {
    return pfc::mulmod32(a_shifted, n, pfc::make_modulus(p));
}
uint64_t t_mulmod64(uint64_t p, uint64_t a_shifted, uint64_t b) {
    return pfc::mulmod64(a_shifted, b, pfc::make_modulus(p));
}
// mulmod64_nofix: *fix set when the result is unreduced.
uint64_t t_mulmod64_nofix(uint64_t p, uint64_t a_shifted, uint64_t b, int* fix) {
    bool f;
    const uint64_t r = pfc::mulmod64_nofix(a_shifted, b, pfc::make_modulus(p), &f);
    *fix = f;
    return r;
}
uint64_t t_sqrmod(uint64_t p, uint64_t a_shifted) { return pfc::sqrmod(a_shifted, pfc::make_modulus(p)); }
uint64_t t_mont_one(uint64_t p) { return pfc::mont_one(pfc::make_modulus(p)); }
int t_modulus_ok(uint64_t p) { return pfc::modulus_ok(p, pfc::make_modulus(p)); }
int t_strong_prp2(uint64_t n) { return pfc::strong_prp2(n); }
int t_is_prime_u32(uint32_t n) { return pfc::is_prime_u32(n); }
// The lockstep test on two values with shared bit and t bounds (the largest
// of the pair, as a wavefront uses them); out[0..1] = verdicts.
void t_strong_prp2_pair(uint64_t n0, uint64_t n1, int out[2]) {
    const uint64_t N[2] = {n0, n1};
    uint32_t bits = 0, tmax = 0;
    for (int k = 0; k < 2; ++k) {
        bits = bits > 64u - (uint32_t)__builtin_clzll(N[k] - 1) ? bits : 64u - (uint32_t)__builtin_clzll(N[k] - 1);
        tmax = tmax > (uint32_t)__builtin_ctzll(N[k] - 1) ? tmax : (uint32_t)__builtin_ctzll(N[k] - 1);
    }
    bool prp[2];
    pfc::strong_prp2_lockstep<2>(N, bits, tmax, prp);
    out[0] = prp[0];
    out[1] = prp[1];
}
uint64_t t_start_residue(const uint64_t* chunks, const uint32_t* groups, uint32_t npos, uint32_t k, uint64_t p) {
    return pfc::start_residue(chunks, groups, npos, k, pfc::make_modulus(p));
}
// The same with the chunks shared over `stride` lanes: the product of the
// partial values of lanes 0 .. stride-1 (the table self-check's split).
uint64_t t_start_residue_split(const uint64_t* chunks, const uint32_t* groups, uint32_t npos, uint32_t k,
                               uint64_t p, uint32_t stride) {
    const pfc::Modulus m = pfc::make_modulus(p);
    uint64_t acc = 1ull << m.s;
    for (uint32_t t = 0; t < stride; ++t)
        acc = pfc::mulmod64(acc, pfc::start_residue(chunks, groups, npos, k, m, t, stride) >> m.s, m);
    return acc;
}
uint64_t t_to_mont(uint64_t p, uint64_t x_shifted) {
    const pfc::Modulus m = pfc::make_modulus(p);
    return pfc::to_mont(x_shifted, pfc::mont_one(m), m);
}
// One walk step as the kernels take it: the 3-by-2 step without its second
// correction (*raw), the attention test on it (*attention), then the
// correction; returns the canonical residue.
uint64_t t_walk_step(uint64_t p, uint64_t a_shifted, uint32_t n, uint64_t* raw, int* attention) {
    const pfc::Modulus m = pfc::make_modulus(p);
    const uint64_t d = pfc::mod_d(m), one = 1ull << m.s;
    bool fix;
    uint64_t x = pfc::mulmod32_nofix(a_shifted, n, pfc::walk_k1(m.d1, m.d0), pfc::walk_k0(m.d0), 0ull - d,
                                     d, m.v, &fix);
    *raw = x;
    uint32_t c, thr;
    pfc::walk_band(d, one, &c, &thr);
    *attention = pfc::walk_attention(x, c, thr);
    return x >= d ? x - d : x;
}
// The walk's test of x against the band B (B = 2^s: +-1).
int t_walk_attention(uint64_t p, uint64_t B, uint64_t x) {
    const pfc::Modulus m = pfc::make_modulus(p);
    uint32_t c, thr;
    pfc::walk_band(pfc::mod_d(m), B, &c, &thr);
    return pfc::walk_attention(x, c, thr);
}
// The narrow walk (pfc::narrow_mul): its constants from the modulus as the
// kernels derive them; out[0] = m, out[1] = np0, out[2] = kp1.
void t_narrow_consts(uint64_t p, uint32_t out[3]) {
    const pfc::Modulus m = pfc::make_modulus(p);
    out[0] = pfc::narrow_m(m.v, m.s);
    out[1] = pfc::narrow_np0(p);
    out[2] = pfc::narrow_kp1(p);
}
int t_narrow_fits(uint64_t p, uint32_t bmax) { return pfc::narrow_fits(p, bmax); }
// `count` narrow steps from x by the multipliers n[0 .. count); returns the
// unreduced result (in [0, 2p) when the step's bounds hold).
uint64_t t_narrow_walk(uint64_t p, uint64_t x, const uint32_t* n, uint32_t count) {
    const pfc::Modulus m = pfc::make_modulus(p);
    const uint32_t mm = pfc::narrow_m(m.v, m.s), np0 = pfc::narrow_np0(p), kp1 = pfc::narrow_kp1(p);
    for (uint32_t i = 0; i < count; ++i) x = pfc::narrow_mul(x, n[i], mm, np0, kp1);
    return x;
}
// The narrow walk's band test of a reduced x against the band B.
int t_narrow_attention(uint64_t p, uint64_t B, uint64_t x) {
    uint64_t c, thr;
    pfc::narrow_band(p, B, &c, &thr);
    return pfc::narrow_attention(x, c, thr);
}
// The medium walk (pfc::medium_mul) at half the shifted scale, D = p << (s - 1):
// `count` steps from x (a multiple of 2^(s-1) below 2D) by n[0 .. count);
// returns the unreduced result.  out[0] = M, out[1] = nd0, out[2] = k1.
void t_medium_consts(uint64_t p, uint32_t out[3]) {
    const pfc::Modulus m = pfc::make_modulus(p);
    const uint64_t D = p << (m.s - 1);
    out[0] = pfc::medium_m(m.v);
    out[1] = pfc::walk_k0((uint32_t)D);
    out[2] = pfc::walk_k1((uint32_t)(D >> 32), (uint32_t)D);
}
int t_medium_fits(uint64_t p, uint32_t bmax) { return pfc::medium_fits(p, bmax); }
uint64_t t_medium_walk(uint64_t p, uint64_t x, const uint32_t* n, uint32_t count) {
    uint32_t c[3];
    t_medium_consts(p, c);
    for (uint32_t i = 0; i < count; ++i) x = pfc::medium_mul(x, n[i], 4u * n[i], c[0], c[1], c[2]);
    return x;
}
// Candidates of [lo, hi) (lo >= 114) exactly as k_generate's lanes compute them.
size_t t_generate(uint64_t lo, uint64_t hi, uint64_t* out, size_t cap) {
    size_t n = 0;
    const uint64_t base30 = lo - lo % 30;
    for (uint64_t start = base30; start < hi && start >= base30; start += 60) {
        uint32_t rmod[27];
        for (int i = 0; i < 27; ++i) rmod[i] = (uint32_t)(start % pfc::kSievePrimes[i]);
        uint64_t c = pfc::window_survivors(start, lo, hi, rmod);
        while (c) {
            const int j = __builtin_ctzll(c);
            c &= c - 1;
            if (pfc::strong_prp2(start + j)) { if (n < cap) out[n] = start + j; ++n; }
        }
    }
    return n;
}
// A work-sized batcher round's division of its updates (pfc::share_steps).
void t_share_steps(const uint64_t* cost, const uint64_t* cap, size_t n, uint64_t budget, uint64_t min_steps,
                   uint64_t* steps) {
    pfc::share_steps(cost, cap, n, budget, min_steps, steps);
}
}
