// Arithmetic shared by the host library and the hipRTC-compiled trajectory kernel.
//
// Everything here is built from integer operations and individually rounded
// IEEE-754 +, -, *, / (no FMA contraction: -ffp-contract=off on both sides),
// so the host and the GPU produce identical bits.
//
//  * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
//    easy as 1, 2, 3", SC'11): counter-based, one stream per trajectory.  It
//    replaces the reference's per-slot curand XORWOW state
//    (src/jit_kernels/simulation.cu:23-32).
//  * mbg_ln: fdlibm-style argument reduction to [sqrt(1/2), sqrt(2)) and the
//    atanh series ln m = 2g(1 + g^2/3 + g^4/5 + ... + g^16/17), g = (m-1)/(m+1),
//    |g| <= 0.1716: truncation below 1e-15 relative, max error about 5 ulp.  Used for the
//    Gillespie waiting time and entropies.
#pragma once

#ifndef MBG_HD
#define MBG_HD inline
#endif

typedef unsigned int mbg_u32;
typedef unsigned long long mbg_u64;

MBG_HD void mbg_philox(mbg_u32 c0, mbg_u32 c1, mbg_u32 c2, mbg_u32 c3, mbg_u32 k0, mbg_u32 k1, mbg_u32* o) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        mbg_u64 p0 = (mbg_u64)0xD2511F53u * c0;
        mbg_u64 p1 = (mbg_u64)0xCD9E8D57u * c2;
        mbg_u32 n0 = (mbg_u32)(p1 >> 32) ^ c1 ^ k0;
        mbg_u32 n2 = (mbg_u32)(p0 >> 32) ^ c3 ^ k1;
        c1 = (mbg_u32)p1;
        c3 = (mbg_u32)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    o[0] = c0;
    o[1] = c1;
    o[2] = c2;
    o[3] = c3;
}

MBG_HD double mbg_ln(double x) {
    mbg_u64 b = __builtin_bit_cast(mbg_u64, x);
    int e = (int)((b >> 52) & 0x7ff);
    if (e == 0) {  // subnormal input
        x = x * 18014398509481984.0;
        b = __builtin_bit_cast(mbg_u64, x);
        e = (int)((b >> 52) & 0x7ff) - 54;
    }
    e -= 1023;
    double m = __builtin_bit_cast(double, (b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e += 1;
    }
    const double f = m - 1.0;
    const double s = m + 1.0;
    const double g = f / s;
    const double z = g * g;
    double p = 1.0 / 17.0;
    p = p * z + 1.0 / 15.0;
    p = p * z + 1.0 / 13.0;
    p = p * z + 1.0 / 11.0;
    p = p * z + 1.0 / 9.0;
    p = p * z + 1.0 / 7.0;
    p = p * z + 1.0 / 5.0;
    p = p * z + 1.0 / 3.0;
    const double r = 2.0 * g;
    const double lnm = r + r * (z * p);
    const double de = (double)e;
    // ln2 split into a head with trailing zero bits and a tail (fdlibm's ln2_hi/ln2_lo)
    return de * 6.93147180369123816490e-01 + (de * 1.90821492927058770002e-10 + lnm);
}

MBG_HD double mbg_log2(double x) { return mbg_ln(x) * 1.4426950408889634; }

// -ln(u) for the uniform u = (v + 1) * 2^-32 in (0, 1].
MBG_HD double mbg_neglog_u32(mbg_u32 v) { return -mbg_ln(((double)v + 1.0) * 2.3283064365386963e-10); }

// FP32 uniform in (0, 1] from 24 bits of a draw (initial states).
MBG_HD float mbg_unit_f32(mbg_u32 v) { return (float)((v >> 8) + 1u) * 5.9604644775390625e-08f; }

// floor(v * total / 2^32) without a 128-bit product: the flip target in [0, total).
MBG_HD mbg_u64 mbg_flip_target(mbg_u32 v, mbg_u64 total) {
    return (mbg_u64)v * (total >> 32) + (((mbg_u64)v * (total & 0xffffffffull)) >> 32);
}
