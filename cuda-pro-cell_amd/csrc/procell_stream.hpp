// procell_stream.hpp -- the random stream of the simulation, device side.
//
// Every draw of a simulation is a pure function of (seed, cell, node,
// attempt): one Philox4x32-10 block (Salmon et al., SC'11) with
//   key     = (seed[31:0], seed[63:32])
//   counter = (node[31:0], node[63:32], cell[31:0], cell[47:32] | attempt << 16)
// node is the heap index of a cell in its division tree (root 1, daughters of
// n are 2n and 2n+1); node 0 carries the root's set-up draws.
//   (cell, 0, 0)    -> the root's draws, one word each: type uniform [0,1)
//                      (word 0), initial phase U (0,1] (word 1), Box-Muller
//                      pair of its first timer's candidates (words 2, 3)
//   (cell, n, 0)    -> Box-Muller pair: timers of daughters 2n and 2n+1
//   (cell, n, a>=1) -> truncated-normal redraw stream of node n; the root's
//                      first timer continues in node 1's stream when both
//                      candidates are <= 0
// This replaces the reference's per-draw cuRAND XORWOW re-initialisation
// (src/utils/util.cu:143-169) and its colliding arithmetic seeds
// (src/simulation/proliferation.cu:330-331, cells_population.cu:108).
//
// The FP64 functions below use only +, -, *, fma and integer/bit operations,
// in a fixed order, plus two lookup tables (procell_math_tables.inc), and the
// library is compiled with -ffp-contract=off, so a host build of the same
// operation sequence reproduces every bit.  Methods:
//   log   : table-driven (Tang, ACM TOMS 16(4), 1990): x = 2^e m, table
//           entry {1/c, -log(1/c)} by the top 7 fraction bits (intervals
//           next to 1 use c = 1), r = fma(m, 1/c, -1), Taylor polynomial of
//           log(1+r) of degree 7, e ln 2 in two parts.
//   sqrt  : Lomont's 64-bit inverse-square-root seed, three Newton steps,
//           one residual correction.
//   sin/cos of 2 pi u : nearest 1/256 turn j (exact reduction), table entry
//           {sin, cos}(2 pi j/256), Taylor polynomials of the remainder,
//           addition theorems.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace procell {
namespace dev {

struct Words {
    uint32_t w0, w1, w2, w3;
};

__device__ __forceinline__ Words philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                               uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        const uint64_t prod_a = (uint64_t)c0 * 0xD2511F53u;
        const uint64_t prod_b = (uint64_t)c2 * 0xCD9E8D57u;
        // three-input XORs in one instruction each (truth table 0x96)
        const uint32_t n0 = __builtin_amdgcn_bitop3_b32((uint32_t)(prod_b >> 32), c1, k0, 0x96);
        const uint32_t n2 = __builtin_amdgcn_bitop3_b32((uint32_t)(prod_a >> 32), c3, k1, 0x96);
        c1 = (uint32_t)prod_b;
        c3 = (uint32_t)prod_a;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return Words{c0, c1, c2, c3};
}

__device__ __forceinline__ Words draw_block(uint64_t seed, uint64_t cell, uint64_t node, uint32_t attempt)
{
    return philox4x32_10((uint32_t)node, (uint32_t)(node >> 32), (uint32_t)cell,
                         ((uint32_t)(cell >> 32) & 0xFFFFu) | (attempt << 16),
                         (uint32_t)seed, (uint32_t)(seed >> 32));
}

// Uniforms from the 53-bit integer m = hi * 2^21 + (lo >> 11):
//   uniform_co = m 2^-53 in [0, 1),  uniform_oc = (m + 1) 2^-53 in (0, 1].
// Each is formed with fmas whose exact results are representable (integers
// below 2^53 scaled by 2^-53), so every step is exact and the value is that
// of the defining formula; the _hi forms are the same values for lo = 0.
__device__ __forceinline__ double uniform_co(uint32_t hi, uint32_t lo)  // [0, 1)
{
    return fma((double)hi, 0x1p-32, (double)(lo >> 11) * 0x1p-53);
}

__device__ __forceinline__ double uniform_oc(uint32_t hi, uint32_t lo)  // (0, 1]
{
    return fma((double)hi, 0x1p-32, fma((double)(lo >> 11), 0x1p-53, 0x1p-53));
}

__device__ __forceinline__ double uniform_co_hi(uint32_t hi) { return (double)hi * 0x1p-32; }

__device__ __forceinline__ double uniform_oc_hi(uint32_t hi) { return fma((double)hi, 0x1p-32, 0x1p-53); }

#define PROCELL_TABLE_QUAL __constant__
#include "procell_math_tables.inc"
#undef PROCELL_TABLE_QUAL

// FP64 coefficients of the stream's log and sin/cos, gathered in one struct so
// that a kernel can load them once into vector registers instead of
// re-materialising every 64-bit literal in scalar registers inside its loop
// (the CU's single scalar unit is shared by all its wavefronts).
struct Coeffs {
    double lp[6];            // log: -1/2, 1/3, -1/4, 1/5, -1/6, 1/7
    double ln2_hi, ln2_lo;   // ln 2 split: hi has 32 significant bits
    double two_pi_hi, two_pi_lo;
    double s1, s2, s3;       // sin t = t + t^3 (s1 + t^2 (s2 + t^2 s3))
    double c1, c2, c3;       // cos t - 1 = t^2 (c1 + t^2 (c2 + t^2 c3))
};

__host__ __device__ constexpr Coeffs coeffs()
{
    return Coeffs{{-0.5, 0x1.5555555555555p-2, -0.25, 0x1.999999999999ap-3, -0x1.5555555555555p-3,
                   0x1.2492492492492p-3},
                  0x1.62e42fee00000p-1,
                  0x1.a39ef35793c76p-33,
                  0x1.921fb54442d18p+2,
                  0x1.1a62633145c07p-52,
                  -0x1.5555555555555p-3,
                  0x1.1111111111111p-7,
                  -0x1.a01a01a01a01ap-13,
                  -0.5,
                  0x1.5555555555555p-5,
                  -0x1.6c16c16c16c17p-10};
}

// Lookup tables of the stream's math; LDS copies in the hot kernels.
struct Tables {
    const double (*log)[2];     // [128] {1/c, -log(1/c)}
    const double (*sincos)[2];  // [256] {sin, cos}(2 pi j/256)
};

// Natural log for positive normal x, in two parts: the table index, then
// the arithmetic given that entry (so a caller can issue the table read
// early and overlap its latency with other work).
__device__ __forceinline__ uint32_t log_index(double x)
{
    return (uint32_t)(__builtin_bit_cast(uint64_t, x) >> 45) & 127u;
}

__device__ __forceinline__ double log_with_entry(double x, uint32_t j, double invc, double logc, const Coeffs &K)
{
    // x = 2^expo m with m in [1, 2), or in [0.75, 1) when the top fraction
    // bit is set (j >= 64): then m is halved exactly by its exponent field
    const uint64_t bits = __builtin_bit_cast(uint64_t, x);
    const uint32_t hw = (uint32_t)(bits >> 32);
    const uint32_t big = j >> 6;
    const int expo = (int)(hw >> 20) - 1023 + (int)big;
    const double m = __builtin_bit_cast(
        double, ((uint64_t)((hw & 0x000FFFFFu) | ((0x3FFu - big) << 20)) << 32) | (uint32_t)bits);
    const double r = fma(m, invc, -1.0);
    const double r2 = r * r;
    double p = K.lp[5];
#pragma unroll
    for (int i = 4; i >= 0; --i)
        p = fma(p, r, K.lp[i]);
    const double de = (double)expo;
    const double hi = fma(de, K.ln2_hi, logc);
    const double lo = de * K.ln2_lo;
    const double t = fma(r2, p, lo);
    return hi + (r + t);
}

__device__ __forceinline__ double log_pos(double x, const Coeffs &K, const Tables &tb)
{
    const uint32_t j = log_index(x);
    return log_with_entry(x, j, tb.log[j][0], tb.log[j][1], K);
}

// Square root for finite a >= 0.
__device__ __forceinline__ double sqrt_nonneg(double a)
{
    const uint64_t ab = __builtin_bit_cast(uint64_t, a);
    double y = __builtin_bit_cast(double, 0x5FE6EB50C7B537A9ull - (ab >> 1));
    const double ha = 0.5 * a;
    double e;
    e = fma(-ha, y * y, 0.5); y = fma(y, e, y);
    e = fma(-ha, y * y, 0.5); y = fma(y, e, y);
    e = fma(-ha, y * y, 0.5); y = fma(y, e, y);
    const double s = a * y;
    const double res = fma(-s, s, a);
    const double out = fma(0.5 * y, res, s);
    return a > 0.0 ? out : 0.0;
}

// sin and cos of 2 pi u with u = uniform_co(hi, lo): the table index (the
// nearest 1/256 turn), then the arithmetic given that entry.
__device__ __forceinline__ uint32_t sincos_index(uint32_t hi)
{
    return ((hi >> 23) + 1u) >> 1;  // 0..256
}

template <bool LoZero = false>
__device__ __forceinline__ void sincos_with_entry(uint32_t hi, uint32_t lo, uint32_t j, double sa, double ca,
                                                  double &sn_out, double &cs_out, const Coeffs &K)
{
    // d = uniform_co(hi, lo) - j 2^-8, exact (|d| <= 1/512, a multiple of
    // 2^-53): the inner fma's exact result (lo >> 11) 2^-53 - j 2^-8 and the
    // outer one's are both representable
    const double d = LoZero ? fma((double)hi, 0x1p-32, (double)j * -0x1p-8)
                            : fma((double)hi, 0x1p-32, fma((double)j, -0x1p-8, (double)(lo >> 11) * 0x1p-53));
    const double t = fma(d, K.two_pi_hi, d * K.two_pi_lo);
    const double t2 = t * t;
    double ps = fma(t2, K.s3, K.s2);
    ps = fma(t2, ps, K.s1);
    const double st = fma(t * t2, ps, t);
    const double cm = t2 * fma(t2, fma(t2, K.c3, K.c2), K.c1);
    sn_out = sa + fma(sa, cm, ca * st);
    cs_out = ca + fma(ca, cm, -(sa * st));
}

__device__ __forceinline__ void sincos_turns(uint32_t hi, uint32_t lo, double &sn_out, double &cs_out,
                                             const Coeffs &K, const Tables &tb)
{
    const uint32_t j = sincos_index(hi);
    sincos_with_entry(hi, lo, j, tb.sincos[j & 255u][0], tb.sincos[j & 255u][1], sn_out, cs_out, K);
}

// Box-Muller (Box & Muller 1958) pair from one block.  Both table entries
// are read before any arithmetic that needs them, so the two LDS round trips
// overlap each other and the uniforms' conversions.
// LoZero: the block's words 1 and 3 are taken as 0 (a root's first timer
// uses words 2 and 3 as (hi, 0) pairs).
template <bool LoZero = false>
__device__ __forceinline__ void normal_pair(const Words &w, double &z0, double &z1, const Coeffs &K,
                                            const Tables &tb)
{
    const double u1 = LoZero ? uniform_oc_hi(w.w0) : uniform_oc(w.w0, w.w1);
    const uint32_t jl = log_index(u1);
    const uint32_t js = sincos_index(w.w2);
    const double invc = tb.log[jl][0], logc = tb.log[jl][1];
    const double sa = tb.sincos[js & 255u][0], ca = tb.sincos[js & 255u][1];
    double sn, cs;
    sincos_with_entry<LoZero>(w.w2, w.w3, js, sa, ca, sn, cs, K);
    const double rho = sqrt_nonneg(-2.0 * log_with_entry(u1, jl, invc, logc, K));
    z0 = rho * cs;
    z1 = rho * sn;
}

// Tables read from constant memory (for kernels off the hot path).
__device__ __forceinline__ Tables global_tables()
{
    return Tables{procell_log_table, procell_sincos_table};
}

}  // namespace dev
}  // namespace procell
