// exp(x) in double-double arithmetic, rounded once to double: the neighbour
// weights' f64::exp (lib/src/ms.rs:1198-1203), which the reference gets from
// the C library. The device's own exp is within about one unit in the last
// place; the weights are cast to f32, so a value near the midpoint of two f32
// neighbours can round to the other one (measured: 1 argument in 2^30).
// Evaluated to about 2^-69 relative, the result is the correctly rounded
// double except within that distance of a double midpoint, and its f32 cast
// matches the C library's but for about one argument in 2^44.
//
// Method (Tang's table-driven exponential, ACM TOMS 15(2), 1989, with
// double-double reduction and reconstruction): n = round(64 x / ln 2),
// j = n mod 64, m = (n - j) / 64; r = x - n ln2/64 as a double-double, with
// ln2/64 split in three parts (the first with 36 significant bits, so n times
// it is exact for |n| < 2^17), |r| <= ln2 / 128; e^r - 1 by its Taylor
// polynomial to r^7 (the tail beyond r in double); exp(x) = 2^m * 2^(j/64) *
// (1 + (e^r - 1)) with the table entry as a double-double. The result is
// within about 2^-69 relative of exp(x) before its single rounding. Table and
// coefficients were computed in 80-digit decimal arithmetic.
#pragma once
#include <math.h>

#include "common.h"

namespace tsg {

// 2^(j/64), j = 0..63, as double-double {hi, lo}: hi = the nearest double,
// lo = the nearest double to the remainder (from 80-digit decimal arithmetic).
constexpr double kExp2Tab[64][2] = {
    {0x1.0000000000000p+0, 0x0.0p+0},
    {0x1.02c9a3e778061p+0, -0x1.19083535b085dp-56},
    {0x1.059b0d3158574p+0, 0x1.d73e2a475b465p-55},
    {0x1.0874518759bc8p+0, 0x1.186be4bb284ffp-57},
    {0x1.0b5586cf9890fp+0, 0x1.8a62e4adc610bp-54},
    {0x1.0e3ec32d3d1a2p+0, 0x1.03a1727c57b53p-59},
    {0x1.11301d0125b51p+0, -0x1.6c51039449b3ap-54},
    {0x1.1429aaea92de0p+0, -0x1.32fbf9af1369ep-54},
    {0x1.172b83c7d517bp+0, -0x1.19041b9d78a76p-55},
    {0x1.1a35beb6fcb75p+0, 0x1.e5b4c7b4968e4p-55},
    {0x1.1d4873168b9aap+0, 0x1.e016e00a2643cp-54},
    {0x1.2063b88628cd6p+0, 0x1.dc775814a8495p-55},
    {0x1.2387a6e756238p+0, 0x1.9b07eb6c70573p-54},
    {0x1.26b4565e27cddp+0, 0x1.2bd339940e9d9p-55},
    {0x1.29e9df51fdee1p+0, 0x1.612e8afad1255p-55},
    {0x1.2d285a6e4030bp+0, 0x1.0024754db41d5p-54},
    {0x1.306fe0a31b715p+0, 0x1.6f46ad23182e4p-55},
    {0x1.33c08b26416ffp+0, 0x1.32721843659a6p-54},
    {0x1.371a7373aa9cbp+0, -0x1.63aeabf42eae2p-54},
    {0x1.3a7db34e59ff7p+0, -0x1.5e436d661f5e3p-56},
    {0x1.3dea64c123422p+0, 0x1.ada0911f09ebcp-55},
    {0x1.4160a21f72e2ap+0, -0x1.ef3691c309278p-58},
    {0x1.44e086061892dp+0, 0x1.89b7a04ef80d0p-59},
    {0x1.486a2b5c13cd0p+0, 0x1.3c1a3b69062f0p-56},
    {0x1.4bfdad5362a27p+0, 0x1.d4397afec42e2p-56},
    {0x1.4f9b2769d2ca7p+0, -0x1.4b309d25957e3p-54},
    {0x1.5342b569d4f82p+0, -0x1.07abe1db13cadp-55},
    {0x1.56f4736b527dap+0, 0x1.9bb2c011d93adp-54},
    {0x1.5ab07dd485429p+0, 0x1.6324c054647adp-54},
    {0x1.5e76f15ad2148p+0, 0x1.ba6f93080e65ep-54},
    {0x1.6247eb03a5585p+0, -0x1.383c17e40b497p-54},
    {0x1.6623882552225p+0, -0x1.bb60987591c34p-54},
    {0x1.6a09e667f3bcdp+0, -0x1.bdd3413b26456p-54},
    {0x1.6dfb23c651a2fp+0, -0x1.bbe3a683c88abp-57},
    {0x1.71f75e8ec5f74p+0, -0x1.16e4786887a99p-55},
    {0x1.75feb564267c9p+0, -0x1.0245957316dd3p-54},
    {0x1.7a11473eb0187p+0, -0x1.41577ee04992fp-55},
    {0x1.7e2f336cf4e62p+0, 0x1.05d02ba15797ep-56},
    {0x1.82589994cce13p+0, -0x1.d4c1dd41532d8p-54},
    {0x1.868d99b4492edp+0, -0x1.fc6f89bd4f6bap-54},
    {0x1.8ace5422aa0dbp+0, 0x1.6e9f156864b27p-54},
    {0x1.8f1ae99157736p+0, 0x1.5cc13a2e3976cp-55},
    {0x1.93737b0cdc5e5p+0, -0x1.75fc781b57ebcp-57},
    {0x1.97d829fde4e50p+0, -0x1.d185b7c1b85d1p-54},
    {0x1.9c49182a3f090p+0, 0x1.c7c46b071f2bep-56},
    {0x1.a0c667b5de565p+0, -0x1.359495d1cd533p-54},
    {0x1.a5503b23e255dp+0, -0x1.d2f6edb8d41e1p-54},
    {0x1.a9e6b5579fdbfp+0, 0x1.0fac90ef7fd31p-54},
    {0x1.ae89f995ad3adp+0, 0x1.7a1cd345dcc81p-54},
    {0x1.b33a2b84f15fbp+0, -0x1.2805e3084d708p-57},
    {0x1.b7f76f2fb5e47p+0, -0x1.5584f7e54ac3bp-56},
    {0x1.bcc1e904bc1d2p+0, 0x1.23dd07a2d9e84p-55},
    {0x1.c199bdd85529cp+0, 0x1.11065895048ddp-55},
    {0x1.c67f12e57d14bp+0, 0x1.2884dff483cadp-54},
    {0x1.cb720dcef9069p+0, 0x1.503cbd1e949dbp-56},
    {0x1.d072d4a07897cp+0, -0x1.cbc3743797a9cp-54},
    {0x1.d5818dcfba487p+0, 0x1.2ed02d75b3707p-55},
    {0x1.da9e603db3285p+0, 0x1.c2300696db532p-54},
    {0x1.dfc97337b9b5fp+0, -0x1.1a5cd4f184b5cp-54},
    {0x1.e502ee78b3ff6p+0, 0x1.39e8980a9cc8fp-55},
    {0x1.ea4afa2a490dap+0, -0x1.e9c23179c2893p-54},
    {0x1.efa1bee615a27p+0, 0x1.dc7f486a4b6b0p-54},
    {0x1.f50765b6e4540p+0, 0x1.9d3e12dd8a18bp-54},
    {0x1.fa7c1819e90d8p+0, 0x1.74853f3a5931ep-55},
};
// 1/k!, k = 1..12, as double-double.
constexpr double kInvFact[12][2] = {
    {0x1.0000000000000p+0, 0x0.0p+0},
    {0x1.0000000000000p-1, 0x0.0p+0},
    {0x1.5555555555555p-3, 0x1.5555555555555p-57},
    {0x1.5555555555555p-5, 0x1.5555555555555p-59},
    {0x1.1111111111111p-7, 0x1.1111111111111p-63},
    {0x1.6c16c16c16c17p-10, -0x1.f49f49f49f49fp-65},
    {0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-73},
    {0x1.a01a01a01a01ap-16, 0x1.a01a01a01a01ap-76},
    {0x1.71de3a556c734p-19, -0x1.c154f8ddc6c00p-73},
    {0x1.27e4fb7789f5cp-22, 0x1.cbbc05b4fa99ap-76},
    {0x1.ae64567f544e4p-26, -0x1.c062e06d1f209p-80},
    {0x1.1eed8eff8d898p-29, -0x1.2aec959e14c06p-83},
};
constexpr double kInvLn2x64 = 0x1.71547652b82fep+6;  // 64 / ln 2
// ln 2 / 64 = kLn2d64a + kLn2d64b + kLn2d64c; kLn2d64a has 36 significant bits
constexpr double kLn2d64a = 0x1.62e42fefa0000p-7, kLn2d64b = 0x1.cf79abc9e3b3ap-46, kLn2d64c = -0x1.ff0342542fc33p-100;

struct DDouble {
    double hi, lo;
};

TSG_HD DDouble dd_two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}

TSG_HD DDouble dd_fast_sum(double a, double b) {  // |a| >= |b|
    const double s = a + b;
    return {s, b - (s - a)};
}

TSG_HD DDouble dd_mul(DDouble a, DDouble b) {
    const double p = a.hi * b.hi;
    const double e = fma(a.hi, b.hi, -p);
    const double cross = a.hi * b.lo + a.lo * b.hi;
    return dd_fast_sum(p, e + cross);
}

TSG_HD DDouble dd_add(DDouble a, DDouble b) {
    const DDouble s = dd_two_sum(a.hi, b.hi);
    return dd_fast_sum(s.hi, s.lo + (a.lo + b.lo));
}

// exp(x) rounded to double; x in [-745, 709] (NaN stays NaN, beyond: 0 or inf).
TSG_HD double exp_dd(double x) {
    if (!(x == x)) return x;
    if (x > 709.0) return HUGE_VAL;
    if (x < -745.0) return 0.0;
    const double n = rint(x * kInvLn2x64);
    const int ni = (int)n;
    const int j = ni & 63;
    const int m = (ni - j) / 64;
    // r = x - n ln2/64 as a double-double: x - n*a is exact (n*a has at most
    // 53 bits and lies within a factor 2 of x, or x itself when n = 0)
    const double t1 = x - n * kLn2d64a;
    const double p2 = n * kLn2d64b;
    const double p2e = fma(n, kLn2d64b, -p2);
    DDouble r = dd_two_sum(t1, -p2);
    r = dd_fast_sum(r.hi, r.lo - p2e - n * kLn2d64c);
    // e^r - 1 = r + r^2 (1/2 + r (1/6 + ... + r / 5040)): the tail beyond the
    // linear term is below 1.5e-5, so double precision leaves it 2^-69 off
    const double h = r.hi;
    double t = kInvFact[6][0];
    for (int k = 5; k >= 1; --k) t = t * h + kInvFact[k][0];
    const DDouble q = dd_fast_sum(h, r.lo + (h * h) * t);
    // 2^(j/64) (1 + q) = T + T q: the product's high part exactly, the rest
    // (all below 2^-52 T) in double, one rounding at the end
    const double th = kExp2Tab[j][0], tl = kExp2Tab[j][1];
    const double yh = th * q.hi;
    const double yl = fma(th, q.hi, -yh);
    const DDouble z = dd_two_sum(th, yh);
    const double rest = z.lo + (yl + (th * q.lo + (tl + tl * q.hi)));
    return ldexp(z.hi + rest, m);
}

}  // namespace tsg
