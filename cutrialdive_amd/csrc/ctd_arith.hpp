// ctd_arith.hpp -- exact modular arithmetic for one odd prime per lane (gfx950).
//
// Replaces the Barrett helpers of modular_arithmetic_detail.hpp:249-424 and
// barrett_reciprocals.hpp:18-40.  The reference derives a 128-bit Barrett
// multiplier per prime with a software 128-bit division; here every
// per-prime set-up (powers, doubling formulas, jump-ahead) runs in 64-bit
// Montgomery form, whose only set-up is a Newton inverse (5 multiplies) and
// one 64-bit remainder.  All results are exact residues, so the factor lists
// are the same as the reference's.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ctd {

__device__ __forceinline__ uint64_t mulhi64(uint64_t a, uint64_t b) { return __umul64hi(a, b); }

// a + b mod p for a, b < p < 2^64 (never overflows).
__device__ __forceinline__ uint64_t add_mod(uint64_t a, uint64_t b, uint64_t p)
{
    uint64_t s = p - b;
    return a >= s ? a - s : a + b;
}
// a - b mod p for a, b < p.
__device__ __forceinline__ uint64_t sub_mod(uint64_t a, uint64_t b, uint64_t p)
{
    return a >= b ? a - b : a + (p - b);
}
__device__ __forceinline__ uint64_t dbl_mod(uint64_t a, uint64_t p) { return add_mod(a, a, p); }

// Montgomery arithmetic with R = 2^64 for any odd p < 2^64.
struct Mont {
    uint64_t p;    // modulus (odd)
    uint64_t pinv; // -p^-1 mod 2^64
    uint64_t one;  // R mod p: the Montgomery form of 1

    __device__ __forceinline__ void init(uint64_t p_)
    {
        p = p_;
        uint64_t inv = p; // correct to 3 bits for odd p
#pragma unroll
        for (int i = 0; i < 5; ++i) inv *= 2 - p * inv; // 6, 12, 24, 48, 96 bits
        pinv = 0 - inv;
        one = (0 - p) % p;
    }

    // a * b * R^-1 mod p for a, b < p.
    __device__ __forceinline__ uint64_t mul(uint64_t a, uint64_t b) const
    {
        uint64_t tlo = a * b, thi = mulhi64(a, b);
        uint64_t q = tlo * pinv;     // t + q p = 0 mod 2^64
        uint64_t qhi = mulhi64(q, p);
        uint64_t u = thi + qhi;
        bool carry = u < thi;
        uint64_t v = u + (tlo != 0); // low words sum to 2^64 unless both are 0
        carry |= v < u;
        if (carry || v >= p) v -= p;
        return v;
    }
    __device__ __forceinline__ uint64_t to_plain(uint64_t a) const { return mul(a, 1); }
    // Montgomery form of a small integer x (x < 2^64), by double-and-add.
    __device__ uint64_t form_of(uint64_t x) const
    {
        uint64_t r = 0;
        if (x == 0) return 0;
        for (int b = 63 - __clzll(x); b >= 0; --b) {
            r = dbl_mod(r, p);
            if ((x >> b) & 1) r = add_mod(r, one, p);
        }
        return r;
    }
    // 2^e in Montgomery form (left-to-right square and double).
    __device__ uint64_t pow2(uint64_t e) const
    {
        uint64_t x = one;
        if (e == 0) return x;
        for (int b = 63 - __clzll(e); b >= 0; --b) {
            x = mul(x, x);
            if ((e >> b) & 1) x = dbl_mod(x, p);
        }
        return x;
    }
    // x^e for x in Montgomery form; result in Montgomery form.
    __device__ uint64_t pow(uint64_t x, uint64_t e) const
    {
        uint64_t r = one;
        while (e) {
            if (e & 1) r = mul(r, x);
            x = mul(x, x);
            e >>= 1;
        }
        return r;
    }
};

} // namespace ctd
