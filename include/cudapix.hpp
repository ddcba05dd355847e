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

// include/V.cuh:6 (DR easy leaves)
inline int64_t V(uint64_t x, uint64_t y)
{
  int64_t o;
  cpx::check(cpx_v_u64(cpx::default_ctx(), x, y, &o));
  return o;
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
  uint64_t base_ = 0, bottom_ = 0, range_ = 0;

 public:
  PiTable() {}
  explicit PiTable(uint64_t range) : range_(range) {}
  PiTable(uint64_t base, uint64_t bottom, uint64_t range = 0)
      : base_(base), bottom_(bottom), range_(range ? range : (base > bottom ? base - bottom : 1)) {}
  ~PiTable()
  {
    if (t_) cpx_pitable_destroy(t_);
  }
  uint64_t get_base() const { return base_; }
  uint64_t get_range() const { return range_; }
  uint64_t get_bottom() const { return bottom_; }
  void set_base(uint64_t b) { base_ = b; reset(); }
  void set_range(uint64_t r) { range_ = r; reset(); }
  void set_bottom(uint64_t b) { bottom_ = b; reset(); }
  uint64_t get_pi_base()
  {
    ensure();
    uint64_t pb = 0;
    cpx_pitable_info(t_, nullptr, nullptr, &pb);
    return pb;
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
    if (!t_) cpx::check(cpx_pitable_create(cpx::default_ctx(), base_, bottom_, range_, &t_));
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

 public:
  explicit Phi(uint32_t max) { cpx::check(cpx_phi_create(cpx::default_ctx(), max, &p_)); }
  Phi(uint32_t max, uint32_t /*p_max*/) : Phi(max) {}
  ~Phi()
  {
    if (p_) cpx_phi_destroy(p_);
  }
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
