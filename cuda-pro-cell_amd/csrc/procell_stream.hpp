// procell_stream.hpp -- the random stream of the simulation, device side.
//
// Every draw of a simulation is a pure function of (seed, cell, node,
// attempt): one Philox4x32-10 block (Salmon et al., SC'11) with
//   key     = (seed[31:0], seed[63:32])
//   counter = (node[31:0], node[63:32], cell[31:0], cell[47:32] | attempt << 16)
// node is the heap index of a cell in its division tree (root 1, daughters of
// n are 2n and 2n+1); node 0 carries the root's set-up draws.
//   (cell, 0, 0)    -> type uniform [0,1) and initial phase U (0,1]
//   (cell, n, 0)    -> Box-Muller pair: timers of daughters 2n and 2n+1
//   (cell, n, a>=1) -> truncated-normal redraw stream of node n; the root's
//                      first timer comes from this stream
// This replaces the reference's per-draw cuRAND XORWOW re-initialisation
// (src/utils/util.cu:143-169) and its colliding arithmetic seeds
// (src/simulation/proliferation.cu:330-331, cells_population.cu:108).
//
// The FP64 functions below use only +, -, *, fma and integer/bit operations,
// in a fixed order, and the library is compiled with -ffp-contract=off, so a
// host build of the same operation sequence reproduces every bit.  Methods:
//   log   : mantissa folded to [sqrt(1/2), sqrt(2)), log(1+f) = 2 atanh(s),
//           s = f/(2+f), fdlibm-style rearrangement, degree-10 Taylor
//           polynomial of atanh in s^2, 1/(2+f) by a quadratic seed and three
//           Newton steps (no division instruction).
//   sqrt  : Lomont's 64-bit inverse-square-root seed, three Newton steps,
//           one residual correction.
//   sin/cos of 2 pi u : exact reduction to the nearest quarter turn, Taylor
//           polynomials of sin(2 pi r) and cos(2 pi r) for |r| <= 1/8.
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

// 53-bit mantissa hi * 2^21 + (lo >> 11), formed exactly.
__device__ __forceinline__ double mantissa53(uint32_t hi, uint32_t lo)
{
    return fma((double)hi, 0x1p21, (double)(lo >> 11));
}

__device__ __forceinline__ double uniform_co(uint32_t hi, uint32_t lo)  // [0, 1)
{
    return mantissa53(hi, lo) * 0x1p-53;
}

__device__ __forceinline__ double uniform_oc(uint32_t hi, uint32_t lo)  // (0, 1]
{
    return (mantissa53(hi, lo) + 1.0) * 0x1p-53;
}

// FP64 coefficients of the stream's log and sin/cos, gathered in one struct so
// that a kernel can load them once into vector registers instead of
// re-materialising every 64-bit literal in scalar registers inside its loop
// (the CU's single scalar unit is shared by all its wavefronts).
struct Coeffs {
    double sqrt2;            // mantissa fold threshold of log
    double rcp[3];           // quadratic seed of 1/(2+f): rcp[0] + rcp[1] d + rcp[2] d^2
    double at[10];           // 2/(2n+1), n = 1..10
    double ln2_hi, ln2_lo;   // ln 2 split: hi has 32 significant bits
    double sn[9];            // (-1)^n (2 pi)^(2n+1) / (2n+1)!, n = 0..8
    double cs[9];            // (-1)^n (2 pi)^(2n) / (2n)!, n = 1..9
};

__host__ __device__ constexpr Coeffs coeffs()
{
    return Coeffs{0x1.6a09e667f3bcdp+0,
                  {0x1.7aa45399395aap+0, -0x1.72711bb58452fp-1, 0x1.df894e1f2af38p-4},
                  {0x1.5555555555555p-1, 0x1.999999999999ap-2, 0x1.2492492492492p-2, 0x1.c71c71c71c71cp-3,
                   0x1.745d1745d1746p-3, 0x1.3b13b13b13b14p-3, 0x1.1111111111111p-3, 0x1.e1e1e1e1e1e1ep-4,
                   0x1.af286bca1af28p-4, 0x1.8618618618618p-4},
                  0x1.62e42fee00000p-1,
                  0x1.a39ef35793c76p-33,
                  {0x1.921fb54442d18p+2, -0x1.4abbce625be53p+5, 0x1.466bc6775aae2p+6, -0x1.32d2cce62bd86p+6,
                   0x1.50783487ee782p+5, -0x1.e3074fde8871fp+3, 0x1.e8f434d018d63p+1, -0x1.6fadb9f155744p-1,
                   0x1.aaec32af93359p-4},
                  {-0x1.3bd3cc9be45dep+4, 0x1.03c1f081b5ac4p+6, -0x1.55d3c7e3cbffap+6, 0x1.e1f506891babbp+5,
                   -0x1.a6d1f2a204a8cp+4, 0x1.f9d38a3763cc3p+2, -0x1.b6e24f44b128fp+0, 0x1.20c62c2f2d7f5p-2,
                   -0x1.2a0c591af8314p-5}};
}

// Natural log for positive normal x.
__device__ __forceinline__ double log_pos(double x, const Coeffs &K)
{
    const uint64_t bits = __builtin_bit_cast(uint64_t, x);
    int expo = (int)(bits >> 52) - 1023;
    double m = __builtin_bit_cast(double, (bits & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    if (m > K.sqrt2) {
        m = m * 0.5;
        expo = expo + 1;
    }
    const double f = m - 1.0;
    const double d = 2.0 + f;
    double r = fma(fma(K.rcp[2], d, K.rcp[1]), d, K.rcp[0]);
    double e;
    e = fma(-d, r, 1.0); r = fma(r, e, r);
    e = fma(-d, r, 1.0); r = fma(r, e, r);
    e = fma(-d, r, 1.0); r = fma(r, e, r);
    const double s = f * r;
    const double z = s * s;
    double p = K.at[9];
#pragma unroll
    for (int i = 8; i >= 0; --i)
        p = fma(p, z, K.at[i]);
    const double R = p * z;
    const double hfsq = (0.5 * f) * f;
    const double dk = (double)expo;
    const double lo_part = fma(s, hfsq + R, dk * K.ln2_lo);
    return dk * K.ln2_hi - ((hfsq - lo_part) - f);
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

// sin and cos of 2 pi u with u = uniform_co(hi, lo).
__device__ __forceinline__ void sincos_turns(uint32_t hi, uint32_t lo, double &sn_out, double &cs_out,
                                             const Coeffs &K)
{
    const double u = uniform_co(hi, lo);
    const uint32_t q = ((hi >> 29) + 1u) >> 1;   // nearest quarter turn, 0..4
    const double r = u - (double)q * 0.25;       // exact, |r| <= 1/8
    const double z = r * r;
    double ps = K.sn[8];
#pragma unroll
    for (int i = 7; i >= 1; --i)
        ps = fma(ps, z, K.sn[i]);
    const double sn = fma(r * z, ps, r * K.sn[0]);
    double pc = K.cs[8];
#pragma unroll
    for (int i = 7; i >= 0; --i)
        pc = fma(pc, z, K.cs[i]);
    const double cs = fma(z, pc, 1.0);
    const uint32_t quad = q & 3u;
    const double a = (quad & 1u) ? cs : sn;   // |sin| candidate
    const double b = (quad & 1u) ? sn : cs;   // |cos| candidate
    sn_out = (quad & 2u) ? -a : a;
    cs_out = ((quad + 1u) & 2u) ? -b : b;
}

// Box-Muller (Box & Muller 1958) pair from one block.
__device__ __forceinline__ void normal_pair(const Words &w, double &z0, double &z1, const Coeffs &K)
{
    const double u1 = uniform_oc(w.w0, w.w1);
    const double rho = sqrt_nonneg(-2.0 * log_pos(u1, K));
    double sn, cs;
    sincos_turns(w.w2, w.w3, sn, cs, K);
    z0 = rho * cs;
    z1 = rho * sn;
}

__device__ __forceinline__ void normal_pair(const Words &w, double &z0, double &z1)
{
    normal_pair(w, z0, z1, coeffs());
}

}  // namespace dev
}  // namespace procell
