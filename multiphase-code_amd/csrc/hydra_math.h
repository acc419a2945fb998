// hydra_math.h -- correctly rounded FP64 division with a shared divisor.
//
// On gfx950 `a / b` compiles to: s = div_scale(b), r = rcp(s), two Newton
// steps r += r*(1 - s*r), q = a'*r, e = a' - s*q, div_fmas(e, r, q),
// div_fixup.  Everything up to the refined reciprocal depends on the divisor
// alone.  When neither operand needs div_scale's exponent scaling and the
// quotient is a normal number, div_scale returns its operand unchanged,
// div_fmas is a plain fma and div_fixup is the identity, so
//     q = a*r;  e = fma(-b, q, a);  a/b == fma(e, r, q)
// bit for bit, with r refined once per divisor.  Operands outside
// [2^-300, 2^300] in magnitude (so their exponents
// differ by less than div_scale's 768 threshold), zero numerators and a non-finite divisor
// take the hardware division, so the result always equals `a / b`.
#ifndef HYDRA_MATH_H
#define HYDRA_MATH_H

#include <hip/hip_runtime.h>

namespace hydra {

struct Divisor {
    double b, r;
    bool ok;
};

// biased exponent in [1023-300, 1023+300]
__device__ __forceinline__ bool exp_mid(double x)
{
    const unsigned hi = (unsigned)(__double_as_longlong(x) >> 52) & 0x7ffu;
    return hi - 723u <= 600u;
}

__device__ __forceinline__ Divisor divisor(double b)
{
    Divisor d;
    d.b = b;
    double r = __builtin_amdgcn_rcp(b);
    double e = fma(-b, r, 1.0);
    r = fma(r, e, r);
    e = fma(-b, r, 1.0);
    r = fma(r, e, r);
    d.r = r;
    d.ok = exp_mid(b);
    return d;
}

__device__ __forceinline__ double divide(double a, const Divisor &d)
{
    const double q = a * d.r;
    const double e = fma(-d.b, q, a);
    double res = fma(e, d.r, q);
    if (__builtin_expect(!(d.ok && exp_mid(a)), 0)) res = a / d.b;
    return res;
}

}  // namespace hydra

#endif
