// Correctly rounded x^y for the FP64 path's two library pow calls, host and
// device: Manning friction's pow(h, 7.0/3.0) (sw_domain_orig.c:2827) and the
// weir's pow(1 - Q2/Q1, 0.385) (sw_domain_orig.c:710).
//
// glibc's pow returns the correctly rounded result in all but extremely rare
// cases, while the GPU math library's pow is only faithful (1 ulp), which
// breaks bitwise parity after a few steps. This evaluates exp(y * ln(x)) in
// double-double arithmetic (relative error below 2^-80 before the final
// rounding), so the rounded result is the correctly rounded one except when
// the exact value lies within 2^-80 of a rounding boundary.
//
// Method (textbook table-driven reductions, e.g. Muller, "Elementary
// Functions", ch. 6):
//   x = 2^e * m, m in [1,2); t = m * RINV[i] - 1 (exact, |t| < 2^-7)
//   ln x = e ln2 - ln RINV[i] + log1p(t)   (cubic part in double-double,
//                                          t^4.. tail in double)
//   y ln x = (128 q + j) ln2/128 + r,
//   x^y = 2^q * 2^(j/128) * exp(r),   |r| <= ln2/256
// Tables: anuga_core_amd/csrc/gen_pow_tables.py. Valid for normal x > 0 whose
// result is a normal double (friction uses x = h >= minimum_allowed_height).
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#define DE_POW_FN __host__ __device__ __forceinline__
#define DE_POW_TABLE static __device__ __host__ const
#else
#define DE_POW_FN static inline
#define DE_POW_TABLE static const
#endif

struct de_dd {
    double hi, lo;
};

DE_POW_FN de_dd de_two_sum(double a, double b)
{
    const double s = a + b;
    const double bb = s - a;
    const double e = (a - (s - bb)) + (b - bb);
    return de_dd{s, e};
}

DE_POW_FN de_dd de_fast_two_sum(double a, double b)
{
    const double s = a + b;
    return de_dd{s, b - (s - a)};
}

DE_POW_FN de_dd de_two_prod(double a, double b)
{
    const double p = a * b;
    return de_dd{p, fma(a, b, -p)};
}

DE_POW_FN de_dd de_dd_add(de_dd a, de_dd b)
{
    de_dd s = de_two_sum(a.hi, b.hi);
    const de_dd t = de_two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = de_fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return de_fast_two_sum(s.hi, s.lo);
}

DE_POW_FN de_dd de_dd_add_d(de_dd a, double b)
{
    de_dd s = de_two_sum(a.hi, b);
    s.lo += a.lo;
    return de_fast_two_sum(s.hi, s.lo);
}

DE_POW_FN de_dd de_dd_mul(de_dd a, de_dd b)
{
    de_dd p = de_two_prod(a.hi, b.hi);
    p.lo += a.hi * b.lo + a.lo * b.hi;
    return de_fast_two_sum(p.hi, p.lo);
}

DE_POW_FN de_dd de_dd_mul_d(de_dd a, double b)
{
    de_dd p = de_two_prod(a.hi, b);
    p.lo += a.lo * b;
    return de_fast_two_sum(p.hi, p.lo);
}

DE_POW_TABLE double kDePowRinv[128] = {
    0x1.fe00000000000p-1, 0x1.fa00000000000p-1, 0x1.f600000000000p-1, 0x1.f200000000000p-1,
    0x1.ee00000000000p-1, 0x1.ea00000000000p-1, 0x1.e800000000000p-1, 0x1.e400000000000p-1,
    0x1.e000000000000p-1, 0x1.dc00000000000p-1, 0x1.da00000000000p-1, 0x1.d600000000000p-1,
    0x1.d200000000000p-1, 0x1.d000000000000p-1, 0x1.cc00000000000p-1, 0x1.c800000000000p-1,
    0x1.c600000000000p-1, 0x1.c200000000000p-1, 0x1.c000000000000p-1, 0x1.bc00000000000p-1,
    0x1.ba00000000000p-1, 0x1.b600000000000p-1, 0x1.b400000000000p-1, 0x1.b000000000000p-1,
    0x1.ae00000000000p-1, 0x1.aa00000000000p-1, 0x1.a800000000000p-1, 0x1.a600000000000p-1,
    0x1.a200000000000p-1, 0x1.a000000000000p-1, 0x1.9e00000000000p-1, 0x1.9a00000000000p-1,
    0x1.9800000000000p-1, 0x1.9600000000000p-1, 0x1.9400000000000p-1, 0x1.9000000000000p-1,
    0x1.8e00000000000p-1, 0x1.8c00000000000p-1, 0x1.8a00000000000p-1, 0x1.8800000000000p-1,
    0x1.8400000000000p-1, 0x1.8200000000000p-1, 0x1.8000000000000p-1, 0x1.7e00000000000p-1,
    0x1.7c00000000000p-1, 0x1.7a00000000000p-1, 0x1.7800000000000p-1, 0x1.7600000000000p-1,
    0x1.7400000000000p-1, 0x1.7200000000000p-1, 0x1.7000000000000p-1, 0x1.6e00000000000p-1,
    0x1.6c00000000000p-1, 0x1.6a00000000000p-1, 0x1.6800000000000p-1, 0x1.6600000000000p-1,
    0x1.6400000000000p-1, 0x1.6200000000000p-1, 0x1.6000000000000p-1, 0x1.5e00000000000p-1,
    0x1.5c00000000000p-1, 0x1.5a00000000000p-1, 0x1.5800000000000p-1, 0x1.5600000000000p-1,
    0x1.5400000000000p-1, 0x1.5200000000000p-1, 0x1.5000000000000p-1, 0x1.5000000000000p-1,
    0x1.4e00000000000p-1, 0x1.4c00000000000p-1, 0x1.4a00000000000p-1, 0x1.4800000000000p-1,
    0x1.4600000000000p-1, 0x1.4600000000000p-1, 0x1.4400000000000p-1, 0x1.4200000000000p-1,
    0x1.4000000000000p-1, 0x1.3e00000000000p-1, 0x1.3e00000000000p-1, 0x1.3c00000000000p-1,
    0x1.3a00000000000p-1, 0x1.3800000000000p-1, 0x1.3800000000000p-1, 0x1.3600000000000p-1,
    0x1.3400000000000p-1, 0x1.3200000000000p-1, 0x1.3200000000000p-1, 0x1.3000000000000p-1,
    0x1.2e00000000000p-1, 0x1.2e00000000000p-1, 0x1.2c00000000000p-1, 0x1.2a00000000000p-1,
    0x1.2a00000000000p-1, 0x1.2800000000000p-1, 0x1.2600000000000p-1, 0x1.2600000000000p-1,
    0x1.2400000000000p-1, 0x1.2200000000000p-1, 0x1.2200000000000p-1, 0x1.2000000000000p-1,
    0x1.1e00000000000p-1, 0x1.1e00000000000p-1, 0x1.1c00000000000p-1, 0x1.1c00000000000p-1,
    0x1.1a00000000000p-1, 0x1.1800000000000p-1, 0x1.1800000000000p-1, 0x1.1600000000000p-1,
    0x1.1600000000000p-1, 0x1.1400000000000p-1, 0x1.1200000000000p-1, 0x1.1200000000000p-1,
    0x1.1000000000000p-1, 0x1.1000000000000p-1, 0x1.0e00000000000p-1, 0x1.0e00000000000p-1,
    0x1.0c00000000000p-1, 0x1.0a00000000000p-1, 0x1.0a00000000000p-1, 0x1.0800000000000p-1,
    0x1.0800000000000p-1, 0x1.0600000000000p-1, 0x1.0600000000000p-1, 0x1.0400000000000p-1,
    0x1.0400000000000p-1, 0x1.0200000000000p-1, 0x1.0200000000000p-1, 0x1.0000000000000p-1,
};

DE_POW_TABLE double kDePowNlogr[256] = {
    0x1.0080559588b35p-8, 0x1.f96638cf63677p-62, 0x1.82448a388a2aap-7, 0x1.04b16137f09a0p-62,
    0x1.432a925980cc1p-6, -0x1.8cdaf39004192p-60, 0x1.c63d2ec14aaf2p-6, -0x1.ce030a686bd86p-60,
    0x1.252f32f8d183fp-5, -0x1.947f792615916p-59, 0x1.67c94f2d4bb58p-5, 0x1.0413e6505e603p-59,
    0x1.894aa149fb343p-5, 0x1.a8be97660a23dp-60, 0x1.ccb73cdddb2ccp-5, -0x1.e48fb0500efd4p-59,
    0x1.08598b59e3a07p-4, -0x1.dd7009902bf32p-58, 0x1.2aa04a44717a5p-4, -0x1.d15d38d2fa3f7p-58,
    0x1.3bdf5a7d1ee64p-4, 0x1.7a976d3b5b45fp-59, 0x1.5e95a4d9791cbp-4, 0x1.f38745c5c450ap-58,
    0x1.8197e2f40e3f0p-4, 0x1.b9f2dffbeed43p-60, 0x1.9335e5d594989p-4, -0x1.478a85704ccb7p-58,
    0x1.b6ac88dad5b1cp-4, -0x1.0057eed1ca59fp-59, 0x1.da727638446a2p-4, 0x1.401fa71733019p-58,
    0x1.ec739830a1120p-4, -0x1.a2bf991780d3fp-59, 0x1.08598b59e3a07p-3, -0x1.dd7009902bf32p-57,
    0x1.1178e8227e47cp-3, -0x1.0e63a5f01c691p-58, 0x1.23d712a49c202p-3, -0x1.6e38161051d69p-57,
    0x1.2d1610c86813ap-3, -0x1.499a3f25af95fp-58, 0x1.3fb45a59928ccp-3, -0x1.d87e6a354d056p-57,
    0x1.4913d8333b561p-3, -0x1.0d5604930f135p-58, 0x1.5bf406b543db2p-3, -0x1.1f5b44c0df7e7p-61,
    0x1.6574ebe8c133ap-3, -0x1.d34f0f4621bedp-60, 0x1.7898d85444c73p-3, 0x1.ef8f6ebcfb201p-58,
    0x1.823c16551a3c2p-3, -0x1.1232ce70be781p-57, 0x1.8beafeb38fe8cp-3, 0x1.55aa8b6997a40p-58,
    0x1.9f6c407089664p-3, 0x1.35a19605e67efp-59, 0x1.a93ed3c8ad9e3p-3, 0x1.bcafa9de97203p-57,
    0x1.b31d8575bce3dp-3, -0x1.6353ab386a94dp-57, 0x1.c6ffbc6f00f71p-3, -0x1.8e58b2c57a4a5p-57,
    0x1.d1037f2655e7bp-3, 0x1.60629242471a2p-57, 0x1.db13db0d48940p-3, 0x1.aa11d49f96cb9p-58,
    0x1.e530effe71012p-3, 0x1.2276041f43042p-59, 0x1.f991c6cb3b379p-3, 0x1.f665066f980a2p-57,
    0x1.01eae5626c691p-2, -0x1.18290bd2932e2p-59, 0x1.07138604d5862p-2, 0x1.cdb16ed4e9138p-56,
    0x1.0c42d676162e3p-2, 0x1.162c79d5d11eep-58, 0x1.1178e8227e47cp-2, -0x1.0e63a5f01c691p-57,
    0x1.1bf99635a6b95p-2, -0x1.12aeb84249223p-57, 0x1.214456d0eb8d4p-2, 0x1.f7ae91aeba60ap-57,
    0x1.269621134db92p-2, 0x1.e0efadd9db02bp-56, 0x1.2bef07cdc9354p-2, -0x1.82dad7fd86088p-56,
    0x1.314f1e1d35ce4p-2, -0x1.3d69909e5c3dcp-56, 0x1.36b6776be1117p-2, -0x1.324f0e883858ep-58,
    0x1.3c25277333184p-2, -0x1.2ad27e50a8ec6p-56, 0x1.419b423d5e8c7p-2, 0x1.0dbb243827392p-57,
    0x1.4718dc271c41bp-2, 0x1.8fb4c14c56eefp-60, 0x1.4c9e09e172c3cp-2, -0x1.123615b147a5dp-58,
    0x1.522ae0738a3d8p-2, -0x1.8f7e9b38a6979p-57, 0x1.57bf753c8d1fbp-2, -0x1.0908d15f88b63p-57,
    0x1.5d5bddf595f30p-2, -0x1.6541148cbb8a2p-56, 0x1.630030b3aac49p-2, 0x1.dc18ce51fff99p-57,
    0x1.68ac83e9c6a14p-2, 0x1.a64eadd740178p-58, 0x1.6e60ee6af1972p-2, 0x1.657c222d868cdp-58,
    0x1.741d876c67bb1p-2, 0x1.84a4ee3059583p-56, 0x1.79e26687cfb3ep-2, -0x1.c168817443f22p-56,
    0x1.7fafa3bd8151cp-2, -0x1.219024acd3b77p-58, 0x1.85855776dcbfbp-2, -0x1.486666443b153p-56,
    0x1.8b639a88b2df5p-2, -0x1.70f2f38238303p-56, 0x1.914a8635bf68ap-2, -0x1.ad4bb98c1f2c5p-56,
    0x1.973a3431356aep-2, -0x1.89d2816cf838fp-57, 0x1.9d32bea15ed3bp-2, 0x1.87bcbcfd3e187p-59,
    0x1.a33440224fa79p-2, -0x1.ba8062860ae23p-57, 0x1.a93ed3c8ad9e3p-2, 0x1.bcafa9de97203p-56,
    0x1.af5295248cdd0p-2, 0x1.9d56c45dd3e86p-56, 0x1.af5295248cdd0p-2, 0x1.9d56c45dd3e86p-56,
    0x1.b56fa04462909p-2, 0x1.494b610665378p-56, 0x1.bb9611b80e2fbp-2, 0x1.6fd02999b21e1p-59,
    0x1.c1c60693fa39ep-2, -0x1.bfc00b8f3feaap-56, 0x1.c7ff9c74554c9p-2, 0x1.223eadb651b4ap-57,
    0x1.ce42f18064743p-2, 0x1.0798270b29f39p-56, 0x1.ce42f18064743p-2, 0x1.0798270b29f39p-56,
    0x1.d490246defa6bp-2, 0x1.d7f4d3b3d406bp-56, 0x1.dae75484c9616p-2, -0x1.0b5837185a661p-56,
    0x1.e148a1a2726cep-2, -0x1.ac81cc8a4dfb8p-56, 0x1.e7b42c3ddad73p-2, 0x1.57d646a17bc6ap-56,
    0x1.e7b42c3ddad73p-2, 0x1.57d646a17bc6ap-56, 0x1.ee2a156b413e5p-2, -0x1.74b71fb5e57e3p-62,
    0x1.f4aa7ee03192dp-2, -0x1.0d487f5aba5e5p-57, 0x1.fb358af7a4884p-2, 0x1.7e8f05924d259p-57,
    0x1.fb358af7a4884p-2, 0x1.7e8f05924d259p-57, 0x1.00e5ae5b207abp-1, 0x1.1713a36138e19p-57,
    0x1.04360be7603adp-1, -0x1.17f9e54e78104p-57, 0x1.078bf0533c568p-1, 0x1.2241edf5fd1f7p-57,
    0x1.078bf0533c568p-1, 0x1.2241edf5fd1f7p-57, 0x1.0ae76e2d054fap-1, 0x1.0d710fcfc4e0dp-55,
    0x1.0e4898611cce1p-1, 0x1.3300f002e836ep-55, 0x1.0e4898611cce1p-1, 0x1.3300f002e836ep-55,
    0x1.11af823c75aa8p-1, -0x1.91eee7772c7c2p-55, 0x1.151c3f6f29612p-1, 0x1.342eb628dba17p-56,
    0x1.151c3f6f29612p-1, 0x1.342eb628dba17p-56, 0x1.188ee40f23ca6p-1, 0x1.89df1568ca0b0p-55,
    0x1.1c07849ae6007p-1, 0x1.59bddae1ccce2p-56, 0x1.1c07849ae6007p-1, 0x1.59bddae1ccce2p-56,
    0x1.1f8635fc61659p-1, -0x1.2164ff40e9817p-56, 0x1.230b0d8bebc98p-1, -0x1.fcc8dbccc25cbp-57,
    0x1.230b0d8bebc98p-1, -0x1.fcc8dbccc25cbp-57, 0x1.269621134db92p-1, 0x1.e0efadd9db02bp-55,
    0x1.2a2786d0ec107p-1, -0x1.6a0c343be95dcp-56, 0x1.2a2786d0ec107p-1, -0x1.6a0c343be95dcp-56,
    0x1.2dbf557b0df43p-1, -0x1.b941ee770436bp-56, 0x1.2dbf557b0df43p-1, -0x1.b941ee770436bp-56,
    0x1.315da4434068bp-1, 0x1.6c3a5f12642c9p-57, 0x1.35028ad9d8c86p-1, -0x1.f01ab6065515cp-56,
    0x1.35028ad9d8c86p-1, -0x1.f01ab6065515cp-56, 0x1.38ae2171976e7p-1, 0x1.21512aa596ea3p-55,
    0x1.38ae2171976e7p-1, 0x1.21512aa596ea3p-55, 0x1.3c6080c36bfb5p-1, 0x1.1930603d87b6ep-56,
    0x1.4019c2125ca93p-1, 0x1.86cf0f38b461ap-57, 0x1.4019c2125ca93p-1, 0x1.86cf0f38b461ap-57,
    0x1.43d9ff2f923c5p-1, -0x1.84f481051f71ap-56, 0x1.43d9ff2f923c5p-1, -0x1.84f481051f71ap-56,
    0x1.47a1527e8a2d3p-1, 0x1.2541aca7d5844p-55, 0x1.47a1527e8a2d3p-1, 0x1.2541aca7d5844p-55,
    0x1.4b6fd6f970c1fp-1, 0x1.c457b531506f6p-55, 0x1.4f45a835a4e19p-1, 0x1.d749362382a77p-56,
    0x1.4f45a835a4e19p-1, 0x1.d749362382a77p-56, 0x1.5322e26867857p-1, 0x1.988ba4aea614dp-56,
    0x1.5322e26867857p-1, 0x1.988ba4aea614dp-56, 0x1.5707a26bb8c66p-1, 0x1.80bff3303dd48p-55,
    0x1.5707a26bb8c66p-1, 0x1.80bff3303dd48p-55, 0x1.5af405c3649e0p-1, -0x1.6714fbcd8135bp-55,
    0x1.5af405c3649e0p-1, -0x1.6714fbcd8135bp-55, 0x1.5ee82aa241920p-1, 0x1.1c066d235ee63p-56,
    0x1.5ee82aa241920p-1, 0x1.1c066d235ee63p-56, 0x1.62e42fefa39efp-1, 0x1.abc9e3b39803fp-56,
};

DE_POW_TABLE double kDePowExp2t[256] = {
    0x1.0000000000000p+0, 0x0.0p+0, 0x1.0163da9fb3335p+0, 0x1.b61299ab8cdb7p-54,
    0x1.02c9a3e778061p+0, -0x1.19083535b085dp-56, 0x1.04315e86e7f85p+0, -0x1.0a31c1977c96ep-54,
    0x1.059b0d3158574p+0, 0x1.d73e2a475b465p-55, 0x1.0706b29ddf6dep+0, -0x1.c91dfe2b13c27p-55,
    0x1.0874518759bc8p+0, 0x1.186be4bb284ffp-57, 0x1.09e3ecac6f383p+0, 0x1.1487818316136p-54,
    0x1.0b5586cf9890fp+0, 0x1.8a62e4adc610bp-54, 0x1.0cc922b7247f7p+0, 0x1.01edc16e24f71p-54,
    0x1.0e3ec32d3d1a2p+0, 0x1.03a1727c57b53p-59, 0x1.0fb66affed31bp+0, -0x1.b9bedc44ebd7bp-57,
    0x1.11301d0125b51p+0, -0x1.6c51039449b3ap-54, 0x1.12abdc06c31ccp+0, -0x1.1b514b36ca5c7p-58,
    0x1.1429aaea92de0p+0, -0x1.32fbf9af1369ep-54, 0x1.15a98c8a58e51p+0, 0x1.2406ab9eeab0ap-55,
    0x1.172b83c7d517bp+0, -0x1.19041b9d78a76p-55, 0x1.18af9388c8deap+0, -0x1.11023d1970f6cp-54,
    0x1.1a35beb6fcb75p+0, 0x1.e5b4c7b4968e4p-55, 0x1.1bbe084045cd4p+0, -0x1.95386352ef607p-54,
    0x1.1d4873168b9aap+0, 0x1.e016e00a2643cp-54, 0x1.1ed5022fcd91dp+0, -0x1.1df98027bb78cp-54,
    0x1.2063b88628cd6p+0, 0x1.dc775814a8495p-55, 0x1.21f49917ddc96p+0, 0x1.2a97e9494a5eep-55,
    0x1.2387a6e756238p+0, 0x1.9b07eb6c70573p-54, 0x1.251ce4fb2a63fp+0, 0x1.ac155bef4f4a4p-55,
    0x1.26b4565e27cddp+0, 0x1.2bd339940e9d9p-55, 0x1.284dfe1f56381p+0, -0x1.a4c3a8c3f0d7ep-54,
    0x1.29e9df51fdee1p+0, 0x1.612e8afad1255p-55, 0x1.2b87fd0dad990p+0, -0x1.10adcd6381aa4p-59,
    0x1.2d285a6e4030bp+0, 0x1.0024754db41d5p-54, 0x1.2ecafa93e2f56p+0, 0x1.1ca0f45d52383p-56,
    0x1.306fe0a31b715p+0, 0x1.6f46ad23182e4p-55, 0x1.32170fc4cd831p+0, 0x1.a9ce78e18047cp-55,
    0x1.33c08b26416ffp+0, 0x1.32721843659a6p-54, 0x1.356c55f929ff1p+0, -0x1.b5cee5c4e4628p-55,
    0x1.371a7373aa9cbp+0, -0x1.63aeabf42eae2p-54, 0x1.38cae6d05d866p+0, -0x1.e958d3c9904bdp-54,
    0x1.3a7db34e59ff7p+0, -0x1.5e436d661f5e3p-56, 0x1.3c32dc313a8e5p+0, -0x1.efff8375d29c3p-54,
    0x1.3dea64c123422p+0, 0x1.ada0911f09ebcp-55, 0x1.3fa4504ac801cp+0, -0x1.7d023f956f9f3p-54,
    0x1.4160a21f72e2ap+0, -0x1.ef3691c309278p-58, 0x1.431f5d950a897p+0, -0x1.1c7dde35f7999p-55,
    0x1.44e086061892dp+0, 0x1.89b7a04ef80d0p-59, 0x1.46a41ed1d0057p+0, 0x1.c944bd1648a76p-54,
    0x1.486a2b5c13cd0p+0, 0x1.3c1a3b69062f0p-56, 0x1.4a32af0d7d3dep+0, 0x1.9cb62f3d1be56p-54,
    0x1.4bfdad5362a27p+0, 0x1.d4397afec42e2p-56, 0x1.4dcb299fddd0dp+0, 0x1.8ecdbbc6a7833p-54,
    0x1.4f9b2769d2ca7p+0, -0x1.4b309d25957e3p-54, 0x1.516daa2cf6642p+0, -0x1.f768569bd93efp-55,
    0x1.5342b569d4f82p+0, -0x1.07abe1db13cadp-55, 0x1.551a4ca5d920fp+0, -0x1.d689cefede59bp-55,
    0x1.56f4736b527dap+0, 0x1.9bb2c011d93adp-54, 0x1.58d12d497c7fdp+0, 0x1.295e15b9a1de8p-55,
    0x1.5ab07dd485429p+0, 0x1.6324c054647adp-54, 0x1.5c9268a5946b7p+0, 0x1.c4b1b816986a2p-60,
    0x1.5e76f15ad2148p+0, 0x1.ba6f93080e65ep-54, 0x1.605e1b976dc09p+0, -0x1.3e2429b56de47p-54,
    0x1.6247eb03a5585p+0, -0x1.383c17e40b497p-54, 0x1.6434634ccc320p+0, -0x1.c483c759d8933p-55,
    0x1.6623882552225p+0, -0x1.bb60987591c34p-54, 0x1.68155d44ca973p+0, 0x1.038ae44f73e65p-57,
    0x1.6a09e667f3bcdp+0, -0x1.bdd3413b26456p-54, 0x1.6c012750bdabfp+0, -0x1.2895667ff0b0dp-56,
    0x1.6dfb23c651a2fp+0, -0x1.bbe3a683c88abp-57, 0x1.6ff7df9519484p+0, -0x1.83c0f25860ef6p-55,
    0x1.71f75e8ec5f74p+0, -0x1.16e4786887a99p-55, 0x1.73f9a48a58174p+0, -0x1.0a8d96c65d53cp-54,
    0x1.75feb564267c9p+0, -0x1.0245957316dd3p-54, 0x1.780694fde5d3fp+0, 0x1.866b80a02162dp-54,
    0x1.7a11473eb0187p+0, -0x1.41577ee04992fp-55, 0x1.7c1ed0130c132p+0, 0x1.f124cd1164dd6p-54,
    0x1.7e2f336cf4e62p+0, 0x1.05d02ba15797ep-56, 0x1.80427543e1a12p+0, -0x1.27c86626d972bp-54,
    0x1.82589994cce13p+0, -0x1.d4c1dd41532d8p-54, 0x1.8471a4623c7adp+0, -0x1.8d684a341cdfbp-55,
    0x1.868d99b4492edp+0, -0x1.fc6f89bd4f6bap-54, 0x1.88ac7d98a6699p+0, 0x1.994c2f37cb53ap-54,
    0x1.8ace5422aa0dbp+0, 0x1.6e9f156864b27p-54, 0x1.8cf3216b5448cp+0, -0x1.0d55e32e9e3aap-56,
    0x1.8f1ae99157736p+0, 0x1.5cc13a2e3976cp-55, 0x1.9145b0b91ffc6p+0, -0x1.dd6792e582524p-54,
    0x1.93737b0cdc5e5p+0, -0x1.75fc781b57ebcp-57, 0x1.95a44cbc8520fp+0, -0x1.64b7c96a5f039p-56,
    0x1.97d829fde4e50p+0, -0x1.d185b7c1b85d1p-54, 0x1.9a0f170ca07bap+0, -0x1.173bd91cee632p-54,
    0x1.9c49182a3f090p+0, 0x1.c7c46b071f2bep-56, 0x1.9e86319e32323p+0, 0x1.824ca78e64c6ep-56,
    0x1.a0c667b5de565p+0, -0x1.359495d1cd533p-54, 0x1.a309bec4a2d33p+0, 0x1.6305c7ddc36abp-54,
    0x1.a5503b23e255dp+0, -0x1.d2f6edb8d41e1p-54, 0x1.a799e1330b358p+0, 0x1.bcb7ecac563c7p-54,
    0x1.a9e6b5579fdbfp+0, 0x1.0fac90ef7fd31p-54, 0x1.ac36bbfd3f37ap+0, -0x1.f9234cae76cd0p-55,
    0x1.ae89f995ad3adp+0, 0x1.7a1cd345dcc81p-54, 0x1.b0e07298db666p+0, -0x1.bdef54c80e425p-54,
    0x1.b33a2b84f15fbp+0, -0x1.2805e3084d708p-57, 0x1.b59728de5593ap+0, -0x1.c71dfbbba6de3p-54,
    0x1.b7f76f2fb5e47p+0, -0x1.5584f7e54ac3bp-56, 0x1.ba5b030a1064ap+0, -0x1.efcd30e54292ep-54,
    0x1.bcc1e904bc1d2p+0, 0x1.23dd07a2d9e84p-55, 0x1.bf2c25bd71e09p+0, -0x1.efdca3f6b9c73p-54,
    0x1.c199bdd85529cp+0, 0x1.11065895048ddp-55, 0x1.c40ab5fffd07ap+0, 0x1.b4537e083c60ap-54,
    0x1.c67f12e57d14bp+0, 0x1.2884dff483cadp-54, 0x1.c8f6d9406e7b5p+0, 0x1.1acbc48805c44p-56,
    0x1.cb720dcef9069p+0, 0x1.503cbd1e949dbp-56, 0x1.cdf0b555dc3fap+0, -0x1.dd83b53829d72p-55,
    0x1.d072d4a07897cp+0, -0x1.cbc3743797a9cp-54, 0x1.d2f87080d89f2p+0, -0x1.d487b719d8578p-54,
    0x1.d5818dcfba487p+0, 0x1.2ed02d75b3707p-55, 0x1.d80e316c98398p+0, -0x1.11ec18beddfe8p-54,
    0x1.da9e603db3285p+0, 0x1.c2300696db532p-54, 0x1.dd321f301b460p+0, 0x1.2da5778f018c3p-54,
    0x1.dfc97337b9b5fp+0, -0x1.1a5cd4f184b5cp-54, 0x1.e264614f5a129p+0, -0x1.7b627817a1496p-54,
    0x1.e502ee78b3ff6p+0, 0x1.39e8980a9cc8fp-55, 0x1.e7a51fbc74c83p+0, 0x1.2d522ca0c8de2p-54,
    0x1.ea4afa2a490dap+0, -0x1.e9c23179c2893p-54, 0x1.ecf482d8e67f1p+0, -0x1.c93f3b411ad8cp-54,
    0x1.efa1bee615a27p+0, 0x1.dc7f486a4b6b0p-54, 0x1.f252b376bba97p+0, 0x1.3a1a5bf0d8e43p-54,
    0x1.f50765b6e4540p+0, 0x1.9d3e12dd8a18bp-54, 0x1.f7bfdad9cbe14p+0, -0x1.dbb12d006350ap-54,
    0x1.fa7c1819e90d8p+0, 0x1.74853f3a5931ep-55, 0x1.fd3c22b8f71f1p+0, 0x1.2eb74966579e7p-57,
};

// x^y for a normal x > 0 whose result is a normal double
DE_POW_FN double de_pow_cr(double x, double y)
{
    const de_dd kLn2{0x1.62e42fefa39efp-1, 0x1.abc9e3b39803fp-56};
    const de_dd kNegLn2_128{-0x1.62e42fefa39efp-8, -0x1.abc9e3b39803fp-63};
    const de_dd kThird{0x1.5555555555555p-2, 0x1.5555555555555p-56};
    const de_dd kSixth{0x1.5555555555555p-3, 0x1.5555555555555p-57};
    const double kInvLn2_128 = 0x1.71547652b82fep+7;   // 128 / ln 2

    // outside the method's domain (zero, negative, subnormal, infinite, NaN:
    // a run that has blown up) the library's pow gives the IEEE special
    // values; the table index below is only valid for normal x > 0
    if (!(x >= 0x1p-1022 && x <= 0x1.fffffffffffffp+1023)) return pow(x, y);

    int e;
    double m = frexp(x, &e);  // m in [0.5, 1)
    m *= 2.0;
    e -= 1;
    const int i = (int)((m - 1.0) * 128.0);
    const double t = fma(m, kDePowRinv[i], -1.0);  // exact

    // log1p(t) = t - t^2/2 + t^3/3 + t^4 * P(t)
    const de_dd t2 = de_two_prod(t, t);
    de_dd l1 = de_two_sum(t, -0.5 * t2.hi);
    l1 = de_dd_add_d(l1, -0.5 * t2.lo);
    const de_dd t3 = de_dd_mul_d(t2, t);
    l1 = de_dd_add(l1, de_dd_mul(t3, kThird));
    const double tail = t2.hi * t2.hi *
        (-0.25 + t * (0.2 + t * (-1.0 / 6 + t * (1.0 / 7 + t * (-0.125 + t * (1.0 / 9 +
         t * (-0.1 + t * (1.0 / 11 + t * (-1.0 / 12 + t * (1.0 / 13))))))))));
    l1 = de_dd_add_d(l1, tail);

    de_dd L = de_dd_mul_d(kLn2, (double)e);
    L = de_dd_add(L, de_dd{kDePowNlogr[2 * i], kDePowNlogr[2 * i + 1]});
    L = de_dd_add(L, l1);

    const de_dd X = de_dd_mul_d(L, y);
    const double kf = rint(X.hi * kInvLn2_128);
    const de_dd r = de_dd_add(X, de_dd_mul_d(kNegLn2_128, kf));
    const int k = (int)kf;
    const int j = k & 127;
    const int q = (k - j) / 128;

    // exp(r) = 1 + r + r^2/2 + r^3/6 + r^4 * Q(r)
    const de_dd r2 = de_dd_mul(r, r);
    de_dd E = de_dd_add_d(r, 1.0);
    E = de_dd_add(E, de_dd{0.5 * r2.hi, 0.5 * r2.lo});
    const de_dd r3 = de_dd_mul(r2, r);
    E = de_dd_add(E, de_dd_mul(r3, kSixth));
    const double rh = r.hi;
    const double etail = r2.hi * r2.hi *
        (1.0 / 24 + rh * (1.0 / 120 + rh * (1.0 / 720 + rh * (1.0 / 5040 + rh * (1.0 / 40320 +
         rh * (1.0 / 362880))))));
    E = de_dd_add_d(E, etail);

    const de_dd R = de_dd_mul(de_dd{kDePowExp2t[2 * j], kDePowExp2t[2 * j + 1]}, E);
    return ldexp(R.hi + R.lo, q);
}

// h^(7/3) of Manning friction (sw_domain_orig.c:2827)
DE_POW_FN double de_pow_7_3(double x) { return de_pow_cr(x, 7.0 / 3.0); }

// x^0.385 of the weir's Villemonte submergence correction
// (sw_domain_orig.c:710), x = 1 - Q2/Q1 in [0, 1]
DE_POW_FN double de_pow_0385(double x) { return x == 0.0 ? 0.0 : de_pow_cr(x, 0.385); }
