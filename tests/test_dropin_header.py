"""CPU test: a reference-style C++ caller compiles against include/cudapix.hpp
(the drop-in for P2.cuh, ordinary.cuh, trivial.cuh, S3.cuh, V.cuh,
pitable.cuh, phi.cuh), links libcudapix_amd.so and runs its host-only self
check.  Device calls are compiled but not executed here; the GPU test in
test_gpu_boundary.py runs the same caller's device branch."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CALLER = r"""
#include "cudapix.hpp"
#include <cmath>
#include <cstdio>
// the reference's call sites, unchanged (src/Deleglise-Rivat/deleglise-rivat.cpp:57-67)
uint64_t dr64(uint64_t x, uint64_t y, uint16_t c) {
  uint64_t p2 = P2(x, y);
  uint64_t ordinary = Ordinary(x, y, c);
  uint64_t trivial = S2_trivial(x, y);
  uint64_t easy = (uint64_t)V(x, y);
  uint64_t hard = S3(x, y, c);
  return ordinary + trivial + easy + hard - 1 - p2;   // + pi(y) by the caller
}
uint128_t dr128(uint128_t x, uint64_t y, uint16_t c) {
  return Ordinary(x, y, c) + S2_trivial(x, y) - P2(x, y);
}
// CUDASieve as the reference calls it (src/P2.cu:100, src/pitable.cu:13, src/S3.cu:27,
// src/sieve/S2_hard_host.cu:65, src/V.cu:64)
uint64_t sieve_calls(uint64_t y) {
  size_t len = 0, len32 = 0, hl = 0;
  uint64_t *d = CudaSieve::getDevicePrimes(0, y, len, 0);
  uint32_t *d32 = CudaSieve::getDevicePrimes32(0, y, len32, 0);
  uint64_t *h = CudaSieve::getHostPrimes(0ull, y, hl, 0);
  uint32_t *bits = CudaSieve::genDeviceBitSieve(0, y, 0);
  uint64_t n = CudaSieve::countPrimes(0, y, 0) + CudaSieve::countPrimes(y);
  cpx_free(cpx::default_ctx(), d);
  cpx_free(cpx::default_ctx(), d32);
  cpx_free(cpx::default_ctx(), bits);
  free(h);
  return n + len + len32 + hl;
}
// the reference's V() body over its PrimeArray overloads (src/V.cu:22-56)
int64_t v_ref(uint64_t x, uint64_t y) {
  int64_t sum = 0, sum_b = 0;
  PrimeArray pi, p, q;
  pi.bottom = 0;
  pi.top = sqrt(x);
  pi.d_primes = CudaSieve::getDevicePrimes(pi.bottom, pi.top, pi.len, 0);
  p.bottom = sqrt(sqrt(x));
  p.top = sqrt(x / y);
  q.bottom = 0;
  q.top = y;
  sum = V_a(x, pi, p, q);
  p.bottom = sqrt(x / y);
  p.top = std::cbrt(x);
  q.bottom = 0;
  q.top = y;
  sum_b = V_b(x, pi, p, q);
  sum += sum_b;
  cpx_free(cpx::default_ctx(), pi.d_primes);
  cpx_free(cpx::default_ctx(), q.d_primes);
  return sum;
}
int main(int argc, char **argv) {
  if (argc > 1) {                       // only with a device
    std::printf("%llu\n", (unsigned long long)sieve_calls(100000));
    PiTable t(1000000, 0, 1000000);
    Phi phi(100000);
    std::printf("%llu %llu\n", (unsigned long long)dr64(1000000000000ull, 20000, 6),
                (unsigned long long)phi.phi(1000, 3));
    (void)t.getCurrent();
    std::printf("%lld %lld\n", (long long)v_ref(1000000000000ull, 20000), (long long)V(1000000000000ull, 20000));
    Phi ph(100000, 1000);
    ph.makeModCache(8);
    std::printf("%u %u %u %llu\n", ph.get_primeListLength(), ph.get_h_primeList()[0],
                ph.get_d_lpf() != nullptr, (unsigned long long)ph.mod_cache()[5]);
    uint64_t *h64 = gen_h_lpf<uint64_t>((uint64_t)1 << 40, ((uint64_t)1 << 40) + 1000);
    int8_t *hmu = gen_h_mu<uint64_t>((uint64_t)1 << 40, ((uint64_t)1 << 40) + 1000);
    std::printf("%llu %d\n", (unsigned long long)h64[0], (int)hmu[0]);
    free(h64);
    free(hmu);
    PiTable t2(1000000, 0, 100000);
    t2.set_pi_base(78498);
    uint32_t *cur = t2.getCurrent();
    std::printf("%llu %d\n", (unsigned long long)t2.get_pi_base(), (int)(cur != nullptr && t2.getMaxRange() > 0));
  }
  std::printf("%d\n", cpx_selftest_arith(1, 10000));
  return 0;
}
"""


def build_caller(tmp_path):
    """Compile CALLER against include/cudapix.hpp and link the library."""
    cxx = shutil.which("g++")
    lib = os.path.join(ROOT, "cudapix_amd", "libcudapix_amd.so")
    assert os.path.exists(lib)
    src = tmp_path / "caller.cpp"
    src.write_text(CALLER)
    exe = tmp_path / "caller"
    cmd = [cxx, "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
           "-L", os.path.join(ROOT, "cudapix_amd"), "-lcudapix_amd",
           "-Wl,-rpath," + os.path.join(ROOT, "cudapix_amd"), "-L/opt/rocm/lib",
           "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.run(cmd, check=True)
    return exe


def test_reference_style_caller_builds_and_links(tmp_path):
    if not shutil.which("g++"):
        pytest.skip("no g++")
    exe = build_caller(tmp_path)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    assert out.strip().splitlines()[-1] == "0"
