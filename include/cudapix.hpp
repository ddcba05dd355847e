// cudapix.hpp -- C++ drop-in for the reference's pi(x) headers, as inline
// forwards to the C ABI in cudapix.h.  A caller of the reference replaces
//   #include "P2.cuh" / "ordinary.cuh" / "trivial.cuh" / "S3.cuh" / "V.cuh" /
//            "pitable.cuh" / "phi.cuh"
// with #include "cudapix.hpp" and keeps its call sites.  Differences that are
// deliberate: results are the exact mathematical term values; errors throw
// cpx::error instead of leaving CUDA state undefined; nothing resets the
// device; uint128_t is this library's own 128-bit type (unsigned __int128).
#pragma once
#include <cstdint>
#include <cstdlib>
#include <new>
#include <stdexcept>
#include <string>
#include "cudapix.h"

typedef unsigned __int128 uint128_t;

namespace cpx {

struct error : std::runtime_error {
  int code;
  error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline cpx_ctx *default_ctx()
{
  static cpx_ctx *c = [] {
    cpx_ctx *h = nullptr;
    int rc = cpx_ctx_create(nullptr, &h);
    if (rc) throw error(rc, cpx_strerror(rc));
    return h;
  }();
  return c;
}

inline void check(int rc)
{
  if (rc) throw error(rc, std::string(cpx_strerror(rc)) + ": " + cpx_ctx_last_error(default_ctx()));
}

inline cpx_u128 to_c(uint128_t v) { return {(uint64_t)v, (uint64_t)(v >> 64)}; }
inline uint128_t from_c(cpx_u128 v) { return ((uint128_t)v.hi << 64) | v.lo; }

}  // namespace cpx

// include/P2.cuh:16-21
inline uint64_t P2(uint64_t x, uint64_t y)
{
  uint64_t o;
  cpx::check(cpx_p2_u64(cpx::default_ctx(), x, y, &o));
  return o;
}
inline uint128_t P2(uint128_t x, uint64_t y)
{
  cpx_u128 o;
  cpx::check(cpx_p2_u128(cpx::default_ctx(), cpx::to_c(x), y, &o));
  return cpx::from_c(o);
}
inline void divXbyY(uint128_t x, uint64_t *d_y, size_t len)
{
  cpx::check(cpx_div_x_by_y(cpx::default_ctx(), cpx::to_c(x), d_y, len));
}

// include/ordinary.cuh:6-7 (the reference returns the signed sum in unsigned types)
inline uint64_t Ordinary(uint64_t x, uint64_t y, uint16_t c)
{
  int64_t o;
  cpx::check(cpx_ordinary_u64(cpx::default_ctx(), x, y, c, &o));
  return (uint64_t)o;
}
inline uint128_t Ordinary(uint128_t x, uint64_t y, uint16_t c)
{
  cpx_i128 o;
  cpx::check(cpx_ordinary_u128(cpx::default_ctx(), cpx::to_c(x), y, c, &o));
  return ((uint128_t)o.hi << 64) | o.lo;
}

// include/trivial.cuh:8-12
inline uint64_t S2_trivial(uint64_t x, uint64_t y)
{
  uint64_t o;
  cpx::check(cpx_s2_trivial_u64(cpx::default_ctx(), x, y, &o));
  return o;
}
inline uint128_t S2_trivial(uint128_t x, uint64_t y)
{
  cpx_u128 o;
  cpx::check(cpx_s2_trivial_u128(cpx::default_ctx(), cpx::to_c(x), y, &o));
  return cpx::from_c(o);
}

// include/S3.cuh:3 (DR hard leaves; signed value returned in the reference's uint64_t)
inline uint64_t S3(uint64_t x, uint64_t y, uint32_t c)
{
  int64_t o;
  cpx::check(cpx_s3_u64(cpx::default_ctx(), x, y, c, &o));
  return (uint64_t)o;
}

// include/cudapix.hpp:14-25 of the reference: the prime windows its V_a / V_b
// take (fields as there; d_primes is freed with cpx_free(cpx::default_ctx(), p)).
struct PrimeArray {
  uint64_t *h_primes = nullptr, *d_primes = nullptr, bottom = 0, top = 0, pi_bottom = 0, pi_top = 0, diff_pi = 0;
  size_t len = 0;
  PrimeArray(uint64_t bottom, uint64_t top) : bottom(bottom), top(top) {}
  PrimeArray() {}
};

// include/V.cuh:6 (DR easy leaves)
inline int64_t V(uint64_t x, uint64_t y)
{
  int64_t o;
  cpx::check(cpx_v_u64(cpx::default_ctx(), x, y, &o));
  return o;
}

namespace cpx {
inline uint64_t isqrt(uint64_t n)
{
  uint64_t r = 0;
  for (int b = 31; b >= 0; b--) {
    const uint64_t t = r | (1ull << b);
    if ((unsigned __int128)t * t <= n) r = t;
  }
  return r;
}
inline uint64_t icbrt(uint64_t n)
{
  uint64_t r = 0;
  for (int b = 21; b >= 0; b--) {
    const uint64_t t = r | (1ull << b);
    if ((unsigned __int128)t * t * t <= n) r = t;
  }
  return r;
}
// V_a / V_b as the reference's V() calls them (src/V.cu:22-56): y = q.top,
// q = [0, y), p = [x^(1/4), sqrt(x/y)) for V_a and [sqrt(x/y), x^(1/3)) for
// V_b.  The arrays are filled as the reference fills them (q's device primes,
// p's prime counts); other windows are refused.
inline int64_t v_half(uint64_t x, PrimeArray &p, PrimeArray &q, bool a)
{
  const uint64_t y = q.top, s = isqrt(x / y);
  const uint64_t lo = a ? isqrt(isqrt(x)) : s, hi = a ? s : icbrt(x);
  if (q.bottom != 0 || p.bottom != lo || p.top != hi)
    throw error(CPX_EINVAL, "V_a/V_b: p must span [x^(1/4), sqrt(x/y)) or [sqrt(x/y), x^(1/3)) and q [0, y)");
  int64_t va = 0, vb = 0;
  check(cpx_v_split_u64(default_ctx(), x, y, &va, &vb));
  if (q.d_primes) cpx_free(default_ctx(), q.d_primes);
  q.d_primes = nullptr;
  check(cpx_get_device_primes(default_ctx(), q.bottom, q.top, &q.d_primes, &q.len));
  check(cpx_count_primes(default_ctx(), 0, p.bottom, &p.pi_bottom));
  check(cpx_count_primes(default_ctx(), 0, p.top, &p.pi_top));
  return a ? va : vb;
}
}  // namespace cpx

// include/V.cuh:7-8
inline int64_t V_a(uint64_t x, PrimeArray &pi, PrimeArray &p, PrimeArray &q)
{
  (void)pi;
  return cpx::v_half(x, p, q, true);
}
inline int64_t V_b(uint64_t x, PrimeArray &pi, PrimeArray &p, PrimeArray &q)
{
  (void)pi;
  return cpx::v_half(x, p, q, false);
}

// include/pitable.cuh:7
inline uint32_t *get_d_piTable(uint32_t hi)
{
  uint32_t *d = nullptr;
  cpx::check(cpx_get_d_pitable(cpx::default_ctx(), hi, &d));
  return d;   // free with cpx_free(cpx::default_ctx(), d)
}

// CUDASieve's calls as the reference makes them (external library; call
// sites src/P2.cu:100,103, src/pitable.cu:13,147, src/S3.cu:27,
// src/sieve/S2_hard_host.cu:65, src/V.cu:64-65): ranges are [bottom, top);
// returned arrays are freed with cpx_free(cpx::default_ctx(), p), host
// arrays with free().
namespace CudaSieve {
inline uint64_t countPrimes(uint64_t bottom, uint64_t top, int device = 0)
{
  (void)device;
  uint64_t n = 0;
  cpx::check(cpx_count_primes(cpx::default_ctx(), bottom, top, &n));
  return n;
}
inline uint64_t countPrimes(uint64_t top) { return countPrimes(0, top); }
inline uint64_t *getDevicePrimes(uint64_t bottom, uint64_t top, size_t &len, int device = 0)
{
  (void)device;
  uint64_t *d = nullptr;
  cpx::check(cpx_get_device_primes(cpx::default_ctx(), bottom, top, &d, &len));
  return d;
}
inline uint32_t *getDevicePrimes32(uint64_t bottom, uint64_t top, size_t &len, int device = 0)
{
  (void)device;
  uint32_t *d = nullptr;
  cpx::check(cpx_get_device_primes32(cpx::default_ctx(), bottom, top, &d, &len));
  return d;
}
inline uint64_t *getHostPrimes(uint64_t bottom, uint64_t top, size_t &len, int device = 0)
{
  (void)device;
  cpx::check(cpx_get_host_primes(cpx::default_ctx(), bottom, top, nullptr, 0, &len));
  uint64_t *h = (uint64_t *)malloc((len ? len : 1) * sizeof(uint64_t));
  if (!h) throw std::bad_alloc();
  size_t n = 0;
  cpx::check(cpx_get_host_primes(cpx::default_ctx(), bottom, top, h, len, &n));
  return h;
}
inline uint32_t *genDeviceBitSieve(uint64_t bottom, uint64_t top, int device = 0)
{
  (void)device;
  uint32_t *d = nullptr;
  size_t nw = 0;
  cpx::check(cpx_gen_device_bit_sieve(cpx::default_ctx(), bottom, top, &d, &nw));
  return d;
}
}  // namespace CudaSieve

// include/pitable.cuh:12-51
class PiTable {
  cpx_pitable *t_ = nullptr;
  uint64_t base_ = 0, bottom_ = 0, range_ = 0, pi_base_ = 0;
  bool pi_set_ = false;

 public:
  PiTable() {}
  explicit PiTable(uint64_t range) : range_(range) { allocate(); }
  PiTable(uint64_t base, uint64_t bottom, uint64_t range = 0)
      : base_(base), bottom_(bottom), range_(range ? range : (base > bottom ? base - bottom : 1)) { allocate(); }
  ~PiTable()
  {
    if (t_) cpx_pitable_destroy(t_);
  }
  uint64_t get_base() const { return base_; }
  uint64_t get_range() const { return range_; }
  uint64_t get_bottom() const { return bottom_; }
  void set_base(uint64_t b) { base_ = b; pi_set_ = false; reset(); }
  void set_range(uint64_t r) { range_ = r; reset(); }
  void set_bottom(uint64_t b) { bottom_ = b; reset(); }
  void set_pi_base(uint64_t pb)
  {
    pi_base_ = pb;
    pi_set_ = true;
    if (t_) cpx::check(cpx_pitable_set_pi_base(t_, pb));
  }
  uint64_t get_pi_base()
  {
    ensure();
    uint64_t pb = 0;
    cpx_pitable_info(t_, nullptr, nullptr, &pb);
    if (!pi_set_ && base_ >= 2 && pb == 0) {          // counted on the first fill
      uint32_t *d = nullptr;
      cpx::check(cpx_pitable_current(t_, &d));
      cpx_pitable_info(t_, nullptr, nullptr, &pb);
    }
    return pb;
  }
  // the device table for the current range (include/pitable.cuh:40-41)
  void allocate() { ensure(); }
  void reallocate()
  {
    if (t_) cpx::check(cpx_pitable_set_range(t_, range_));
    else ensure();
  }
  uint64_t getMaxRange()
  {
    uint64_t r = 0;
    cpx::check(cpx_pitable_max_range(cpx::default_ctx(), &r));
    return r;
  }
  uint64_t setMaxRange()
  {
    range_ = getMaxRange();
    reallocate();
    return range_;
  }
  uint32_t *getCurrent() { return step(cpx_pitable_current); }
  uint32_t *getNextUp() { return step(cpx_pitable_next_up); }
  uint32_t *getNextDown() { return step(cpx_pitable_next_down); }
  uint64_t getNextBaseUp() const { return base_ + range_; }
  uint64_t getNextBaseDown() const { return base_ > bottom_ + range_ ? base_ - range_ : bottom_; }

 private:
  void reset()
  {
    if (t_) cpx_pitable_destroy(t_);
    t_ = nullptr;
  }
  void ensure()
  {
    if (t_) return;
    cpx::check(cpx_pitable_create(cpx::default_ctx(), base_, bottom_, range_, &t_));
    if (pi_set_) cpx::check(cpx_pitable_set_pi_base(t_, pi_base_));
  }
  uint32_t *step(int (*f)(cpx_pitable *, uint32_t **))
  {
    ensure();
    uint32_t *d = nullptr;
    cpx::check(f(t_, &d));
    cpx_pitable_info(t_, &base_, &range_, nullptr);
    return d;
  }
};

// include/phi.cuh:17-36
class Phi {
  cpx_phi *p_ = nullptr;
  uint64_t *cache_ = nullptr;   // makeModCache: cache_[k-1] = phi(p_1...p_k, k)
  uint32_t c_ = 0;

 public:
  explicit Phi(uint32_t max) { cpx::check(cpx_phi_create(cpx::default_ctx(), max, &p_)); }
  Phi(uint32_t max, uint32_t p_max) { cpx::check(cpx_phi_create_pmax(cpx::default_ctx(), max, p_max, &p_)); }
  ~Phi()
  {
    if (p_) cpx_phi_destroy(p_);
    free(cache_);
  }
  Phi(const Phi &) = delete;
  Phi &operator=(const Phi &) = delete;
  // include/phi.cuh:23-26
  uint32_t *get_d_lpf()
  {
    uint32_t *d = nullptr;
    cpx::check(cpx_phi_lpf(p_, &d, nullptr));
    return d;
  }
  uint32_t *get_d_primeList()
  {
    uint32_t *d = nullptr;
    cpx::check(cpx_phi_prime_list(p_, &d, nullptr, nullptr));
    return d;
  }
  uint32_t *get_h_primeList()
  {
    const uint32_t *h = nullptr;
    cpx::check(cpx_phi_prime_list(p_, nullptr, &h, nullptr));
    return const_cast<uint32_t *>(h);
  }
  uint32_t get_primeListLength()
  {
    uint32_t n = 0;
    cpx::check(cpx_phi_prime_list(p_, nullptr, nullptr, &n));
    return n;
  }
  // include/phi.cuh:32
  void makeModCache(uint32_t pi_max)
  {
    uint64_t *c = (uint64_t *)malloc((pi_max > 1 ? pi_max - 1 : 1) * sizeof(uint64_t));
    if (!c) throw std::bad_alloc();
    const int rc = cpx_phi_mod_cache(p_, pi_max, c);
    if (rc) { free(c); cpx::check(rc); }
    free(cache_);
    cache_ = c;
    c_ = pi_max;
  }
  const uint64_t *mod_cache() const { return cache_; }
  uint64_t phi(uint32_t x, uint32_t a)
  {
    uint64_t o;
    cpx::check(cpx_phi_eval(p_, x, a, &o));
    return o;
  }
  uint32_t *generateRange(uint32_t num, uint32_t a)
  {
    uint32_t *d = nullptr;
    cpx::check(cpx_phi_range(p_, num, a, &d));
    return d;   // free with cpx_free(cpx::default_ctx(), d)
  }
};

// include/sieve/lpf_mu.cuh:14-21: odd-only windows n = bottom + 2i + 1 < top
// (bottom even), freed with cpx_free(cpx::default_ctx(), p); the host
// versions return malloc'd copies (free()).
template <typename T>
T *gen_d_lpf(T bottom, T top);
template <>
inline uint32_t *gen_d_lpf<uint32_t>(uint32_t bottom, uint32_t top)
{
  uint32_t *d = nullptr;
  cpx::check(cpx_gen_d_lpf(cpx::default_ctx(), bottom, top, &d));
  return d;
}
template <>
inline uint64_t *gen_d_lpf<uint64_t>(uint64_t bottom, uint64_t top)
{
  uint64_t *d = nullptr;
  cpx::check(cpx_gen_d_lpf64(cpx::default_ctx(), bottom, top, &d));
  return d;
}
template <typename T>
int8_t *gen_d_mu(T bottom, T top)
{
  int8_t *d = nullptr;
  cpx::check(cpx_gen_d_mu(cpx::default_ctx(), bottom, top, &d));
  return d;
}
template <typename T>
T *gen_h_lpf(T bottom, T top)
{
  const size_t n = (size_t)(top - bottom) / 2;
  T *d = gen_d_lpf<T>(bottom, top);
  T *h = (T *)malloc((n ? n : 1) * sizeof(T));
  const int rc = h ? cpx_copy_to_host(cpx::default_ctx(), h, d, n * sizeof(T)) : CPX_ENOMEM;
  cpx_free(cpx::default_ctx(), d);
  if (rc) { free(h); cpx::check(rc); }
  return h;
}
template <typename T>
int8_t *gen_h_mu(T bottom, T top)
{
  const size_t n = (size_t)(top - bottom) / 2;
  int8_t *d = gen_d_mu<T>(bottom, top);
  int8_t *h = (int8_t *)malloc(n ? n : 1);
  const int rc = h ? cpx_copy_to_host(cpx::default_ctx(), h, d, n) : CPX_ENOMEM;
  cpx_free(cpx::default_ctx(), d);
  if (rc) { free(h); cpx::check(rc); }
  return h;
}

// src/Deleglise-Rivat/deleglise-rivat.cpp:50 / :124
namespace deleglise_rivat {
inline uint64_t pi(uint64_t x, uint64_t y = 0, uint16_t c = 6)
{
  uint64_t o;
  cpx::check(cpx_pi_u64(cpx::default_ctx(), x, y, c, &o, nullptr));
  return o;
}
inline uint128_t pi(uint128_t x, uint64_t y = 0, uint16_t c = 6)
{
  cpx_u128 o;
  cpx::check(cpx_pi_u128(cpx::default_ctx(), cpx::to_c(x), y, c, &o, nullptr));
  return cpx::from_c(o);
}
}  // namespace deleglise_rivat
