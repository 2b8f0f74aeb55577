// rg_cos.h -- sin x and cos x for 0 <= x < 2 pi with no lane branches: the angle 2 pi u2 of the Box-Muller
// transform of the disorder generator.  The library's cos carries a large-argument reduction behind a
// branch that never runs for this range but splits the code into basic blocks, so that the generator
// cannot be interleaved with matrix-core products.  RG_HD defaults to __host__ __device__; a host-only
// build (the accuracy test) defines it empty before including this file.
//
// Reduction (Cody and Waite, with pi/2 split into three doubles C0 + C1 + C2, each the rounding of the
// remainder of the previous ones): k = rint(x 2/pi) in 0..4, t0 = x - k C0 is exact (x and k C0 are
// multiples of 2^-53 and |t0| < 1: one rounding of the exact FMA result loses nothing), k C1 = p + pe
// exactly (FMA), t0 - p = s + e exactly (two-sum), and r = rh + rl = s + (e - pe - k C2).  The smallest |r|
// over doubles below 2 pi is 6.1e-17 (x = RN(pi/2)); the rounding of k C2 and the omitted remainder of pi/2
// are below 2^-159.  Kernels on |rh| <= pi/4 + 2^-48: the Taylor series of sin to r^17 and of cos to r^18
// (first omitted terms below 2^-62 relative), with the first-order terms of rl and the exact error of rh^2.  Then
// (sin x, cos x) from (sin r, cos r) by the quadrant k mod 4, selected without branching.
#pragma once

#ifndef RG_HD
#define RG_HD __host__ __device__
#endif

RG_HD inline void rg_sincos_0_2pi(double x, double *sinx, double *cosx)
{
#pragma clang fp contract(off)  // the two-sums must round each operation as written
    const double C0 = 0x1.921fb54442d18p+0, C1 = 0x1.1a62633145c07p-54, C2 = -0x1.f1976b7ed8fbcp-110;
    const double k = rint(x * 0x1.45f306dc9c883p-1);
    const double t0 = fma(-k, C0, x);
    const double p = k * C1, pe = fma(k, C1, -p);
    const double s = t0 - p, bb = s - t0;
    const double e = (t0 - (s - bb)) + (-p - bb);
    const double lo = (e - pe) - k * C2;
    const double rh = s + lo, rl = lo - (rh - s);

    const double z = rh * rh, ze = fma(rh, rh, -z);  // rh^2 = z + ze
    // sin r = rh + rh z S(z) + rl (1 - z/2), S(z) = sum_{n>=1} (-1)^n z^(n-1) / (2n+1)!
    double sp = 0x1.952c77030ad4ap-49;           // 1/17!
    sp = fma(sp, z, -0x1.ae7f3e733b81fp-41);    // -1/15!
    sp = fma(sp, z, 0x1.6124613a86d09p-33);     // 1/13!
    sp = fma(sp, z, -0x1.ae64567f544e4p-26);    // -1/11!
    sp = fma(sp, z, 0x1.71de3a556c734p-19);     // 1/9!
    sp = fma(sp, z, -0x1.a01a01a01a01ap-13);    // -1/7!
    sp = fma(sp, z, 0x1.1111111111111p-7);      // 1/5!
    sp = fma(sp, z, -0x1.5555555555555p-3);     // -1/3!
    const double sinr = rh + fma(rh * z, sp, rl * fma(-0.5, z, 1.0));
    // cos r = 1 - z/2 + z^2 C(z) - rh rl - ze/2, C(z) = sum_{n>=2} (-1)^n z^(n-2) / (2n)!
    double cp = -0x1.6827863b97d97p-53;          // -1/18!
    cp = fma(cp, z, 0x1.ae7f3e733b81fp-45);      // 1/16!
    cp = fma(cp, z, -0x1.93974a8c07c9dp-37);     // -1/14!
    cp = fma(cp, z, 0x1.1eed8eff8d898p-29);      // 1/12!
    cp = fma(cp, z, -0x1.27e4fb7789f5cp-22);     // -1/10!
    cp = fma(cp, z, 0x1.a01a01a01a01ap-16);      // 1/8!
    cp = fma(cp, z, -0x1.6c16c16c16c17p-10);     // -1/6!
    cp = fma(cp, z, 0x1.5555555555555p-5);       // 1/4!
    const double hz = 0.5 * z, w = 1.0 - hz;
    const double tail = ((1.0 - w) - hz) + (fma(z * z, cp, -rh * rl) - 0.5 * ze);
    const double cosr = w + tail;

    // x = r + q pi/2: (sin x, cos x) = (sin r, cos r), (cos r, -sin r), (-sin r, -cos r), (-cos r, sin r)
    const int q = (int)k & 3;
    const double a = (q & 1) ? cosr : sinr, c = (q & 1) ? sinr : cosr;
    *sinx = (q == 2 || q == 3) ? -a : a;
    *cosx = (q == 1 || q == 2) ? -c : c;
}
