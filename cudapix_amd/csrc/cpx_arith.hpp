// cpx_arith.hpp -- exact integer arithmetic shared by host and device code.
//
// Replaces the external CUDA-uint128 helpers the reference calls
// (_isqrt/_icbrt/_iqrt, div128to64, div128to128, mul128; e.g.
// src/Deleglise-Rivat/deleglise-rivat.cpp:96-103, src/ordinary.cu:263-265,
// src/P2.cu:22-23) with exact algorithms written from the mathematics:
//   * integer k-th roots: floating estimate, then exact correction steps;
//   * 128/64 division: normalised two-digit schoolbook division with 32-bit
//     half-digits (Knuth, TAOCP vol. 2, 4.3.1 Algorithm D, specialised to a
//     two-word dividend and one-word divisor);
//   * small-quotient 64-bit division: an f64 quotient estimate, exact after
//     one signed remainder correction (valid when the quotient is < 2^50).
// Nothing here copies a CUDA math routine; every function is checked against
// native __int128 arithmetic on the host (cpx_selftest_arith).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define CPX_HD __host__ __device__ __forceinline__
#else
#define CPX_HD inline
#endif

typedef unsigned __int128 u128;
typedef __int128 i128;

CPX_HD int cpx_clz64(uint64_t v) { return v ? __builtin_clzll(v) : 64; }

// floor((u1 * 2^64 + u0) / v), requires u1 < v.  *r = remainder.
CPX_HD uint64_t cpx_divlu(uint64_t u1, uint64_t u0, uint64_t v, uint64_t *r)
{
  const uint64_t b = 1ull << 32;
  int s = cpx_clz64(v);
  v <<= s;                                   // normalise: top bit of v set
  uint64_t vn1 = v >> 32, vn0 = v & 0xffffffffull;
  uint64_t un32 = s ? (u1 << s) | (u0 >> (64 - s)) : u1;
  uint64_t un10 = u0 << s;
  uint64_t un1 = un10 >> 32, un0 = un10 & 0xffffffffull;

  uint64_t q1 = un32 / vn1, rhat = un32 - q1 * vn1;
  while (q1 >= b || q1 * vn0 > ((rhat << 32) | un1)) {
    q1--; rhat += vn1;
    if (rhat >= b) break;
  }
  uint64_t un21 = (un32 << 32) + un1 - q1 * v;   // exact mod 2^64, < v

  uint64_t q0 = un21 / vn1;
  rhat = un21 - q0 * vn1;
  while (q0 >= b || q0 * vn0 > ((rhat << 32) | un0)) {
    q0--; rhat += vn1;
    if (rhat >= b) break;
  }
  if (r) *r = ((un21 << 32) + un0 - q0 * v) >> s;
  return (q1 << 32) | q0;
}

// 128-bit by 64-bit division -> 128-bit quotient (div128to128 replacement).
CPX_HD u128 cpx_div128_64(u128 a, uint64_t d, uint64_t *rem)
{
  uint64_t hi = (uint64_t)(a >> 64), lo = (uint64_t)a;
  if (hi == 0) {
    if (rem) *rem = lo % d;
    return lo / d;
  }
  uint64_t qh = hi / d, r1 = hi - qh * d;       // r1 < d
  uint64_t r;
  uint64_t ql = cpx_divlu(r1, lo, d, &r);
  if (rem) *rem = r;
  return ((u128)qh << 64) | ql;
}

// floor(a / b) when the quotient is known to be < 2^50 and b < 2^53.
CPX_HD uint64_t cpx_div_smallq(uint64_t a, uint64_t b)
{
  uint64_t q = (uint64_t)((double)a / (double)b);
  int64_t r = (int64_t)(a - q * b);             // true remainder in [-b, 2b)
  if (r < 0) q--;
  else if ((uint64_t)r >= b) q++;
  return q;
}

// a mod p for a < 2^64, p < 2^32 (quotient may be large: uses the exact
// two-step path when a / p could reach 2^50).
CPX_HD uint64_t cpx_mod64_32(uint64_t a, uint32_t p)
{
  if (a < (1ull << 50)) {
    uint64_t q = cpx_div_smallq(a, p);
    return a - q * p;
  }
  return a % p;
}

// floor(x / d) for a 128-bit x whose quotient by d fits 64 bits.
CPX_HD uint64_t cpx_div128_to64(u128 x, uint64_t d)
{
  uint64_t hi = (uint64_t)(x >> 64), lo = (uint64_t)x;
  if (hi == 0) {
    if (d < (1ull << 53) && (u128)lo < ((u128)d << 50)) return cpx_div_smallq(lo, d);
    return lo / d;
  }
  uint64_t r;
  return cpx_divlu(hi, lo, d, &r);               // requires hi < d (quotient < 2^64)
}

// ------------------------------------------------------------- integer roots
// (host only: query bounds are planned on the host)

inline uint64_t cpx_isqrt128(u128 x)
{
  if (x == 0) return 0;
  double d = (double)x;
  uint64_t r;
  // sqrt via the host/device libm call; exactness comes from the fix-ups.
  double s = __builtin_sqrt(d);
  r = s >= 18446744073709551615.0 ? 0xffffffffffffffffull : (uint64_t)s;
  while ((u128)r * r > x) r--;
  while (r < 0xffffffffffffffffull && (u128)(r + 1) * (r + 1) <= x) r++;
  return r;
}

inline uint64_t cpx_icbrt128(u128 x)
{
  if (x == 0) return 0;
  double c = __builtin_cbrt((double)x);
  uint64_t r = (uint64_t)c;
  while (r > 0 && (u128)r * r * r > x) r--;
  while ((u128)(r + 1) * (r + 1) * (r + 1) <= x) r++;
  return r;
}

inline uint64_t cpx_iroot4_128(u128 x) { return cpx_isqrt128((u128)cpx_isqrt128(x)); }

// -------------------------------------------------------------- popcount

CPX_HD uint32_t cpx_popc32(uint32_t v) { return (uint32_t)__builtin_popcount(v); }
CPX_HD uint32_t cpx_popc64(uint64_t v) { return (uint32_t)__builtin_popcountll(v); }

// mask of the low k bits of a 32-bit word, k in [0, 32]
CPX_HD uint32_t cpx_lowmask32(uint32_t k) { return (uint32_t)((1ull << k) - 1ull); }
