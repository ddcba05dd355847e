"""ctypes wrapper of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, always as the values GPU results are compared with or
the timed CPU baseline,
never by the shipped library.  See oracle.c for the reference file:line each
term restates.
"""
import ctypes
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")
_lib = None

U64 = ctypes.c_uint64
MASK64 = (1 << 64) - 1


def build():
    """Compile oracle.c into liboracle.so (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "liboracle.so"], check=True)


def use_native():
    """Rebuild the oracle with -O3 -march=native on this machine and route
    every later call through that build (for timing the CPU baseline; the
    portable build stays the one the tests load).  Returns the compile flags."""
    global _lib
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "liboracle_native.so"], check=True)
    _lib = _load(os.path.join(_HERE, "liboracle_native.so"))
    return "-O3 -march=native -fopenmp"


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        _lib = _load(_SO)
    return _lib


def _load(path):
    L = ctypes.CDLL(path)
    for name in ("orc_isqrt", "orc_icbrt", "orc_iroot4"):
        getattr(L, name).restype = U64
        getattr(L, name).argtypes = [U64, U64]
    L.orc_count_primes.restype = U64
    L.orc_count_primes.argtypes = [U64, U64, ctypes.c_int]
    L.orc_primes.restype = ctypes.c_size_t
    L.orc_primes.argtypes = [U64, ctypes.c_void_p, ctypes.c_size_t]
    L.orc_mertens.restype = ctypes.c_int64
    L.orc_mertens.argtypes = [U64]
    L.orc_phi_brute.restype = U64
    L.orc_phi_brute.argtypes = [U64, ctypes.c_uint32]
    L.orc_phi_legendre.restype = U64
    L.orc_phi_legendre.argtypes = [U64, ctypes.c_uint32]
    L.orc_mu_lpf_odd.restype = None
    L.orc_mu_lpf_odd.argtypes = [U64, ctypes.c_void_p, ctypes.c_void_p]
    L.orc_mu_lpf_window.restype = None
    L.orc_mu_lpf_window.argtypes = [U64, U64, ctypes.c_void_p, ctypes.c_void_p]
    L.orc_choose_y.restype = U64
    L.orc_choose_y.argtypes = [U64, U64, ctypes.c_double]
    L.orc_clamp_c.restype = ctypes.c_uint32
    L.orc_clamp_c.argtypes = [U64, U64, ctypes.c_uint32]
    out2 = ctypes.POINTER(U64)
    L.orc_p2.argtypes = [U64, U64, U64, ctypes.c_int, out2]
    L.orc_ordinary.argtypes = [U64, U64, U64, ctypes.c_uint32, ctypes.c_int, out2]
    L.orc_trivial.argtypes = [U64, U64, U64, ctypes.c_uint32, out2]
    L.orc_easy.argtypes = [U64, U64, U64, ctypes.c_uint32, ctypes.c_int, out2]
    L.orc_easy_part.argtypes = [U64, U64, U64, ctypes.c_uint32, ctypes.c_int, ctypes.c_int, out2]
    L.orc_hard.argtypes = [U64, U64, U64, ctypes.c_uint32, ctypes.c_int, out2]
    L.orc_pi_dr.argtypes = [U64, U64, U64, ctypes.c_uint32, ctypes.c_int, out2]
    L.orc_sigma.argtypes = [U64, U64, U64, ctypes.c_int, out2]
    L.orc_sigma.restype = None
    for n in ("orc_p2", "orc_ordinary", "orc_trivial", "orc_easy", "orc_easy_part", "orc_hard",
              "orc_pi_dr"):
        getattr(L, n).restype = None
    return L


def _split(x):
    return (x >> 64) & MASK64, x & MASK64


def _s128(lo, hi):
    v = (hi << 64) | lo
    return v - (1 << 128) if v >> 127 else v


def isqrt(x):
    return lib().orc_isqrt(*_split(x))


def icbrt(x):
    return lib().orc_icbrt(*_split(x))


def iroot4(x):
    return lib().orc_iroot4(*_split(x))


def choose_y(x, alpha):
    return lib().orc_choose_y(*_split(x), float(alpha))


def clamp_c(x, c):
    return lib().orc_clamp_c(*_split(x), c)


def count_primes(lo, hi, threads=0):
    """Number of primes in [lo, hi) by a segmented sieve of Eratosthenes."""
    return lib().orc_count_primes(lo, hi, threads)


def primes(limit):
    n = lib().orc_primes(limit, None, 0)
    buf = (ctypes.c_uint32 * max(n, 1))()
    lib().orc_primes(limit, buf, n)
    return list(buf[:n])


def mertens(n):
    return lib().orc_mertens(n)


def phi_brute(n, b):
    return lib().orc_phi_brute(n, b)


def phi_legendre(x, a):
    return lib().orc_phi_legendre(x, a)


def mu_lpf_odd(y):
    """(mu, lpf) lists for odd n = 2i+1 <= y (lpf(1) = 2^64-1)."""
    n = (y + 1) // 2
    mu = (ctypes.c_int8 * max(n, 1))()
    lpf = (U64 * max(n, 1))()
    lib().orc_mu_lpf_odd(y, mu, lpf)
    return list(mu[:n]), list(lpf[:n])


def mu_lpf_window(bottom, top):
    """(mu, lpf) lists for the odd n = bottom + 2i + 1 < top (bottom even), by
    trial division of each n (lpf(1) = 2^64-1)."""
    n = (top - bottom) // 2
    mu = (ctypes.c_int8 * max(n, 1))()
    lpf = (U64 * max(n, 1))()
    lib().orc_mu_lpf_window(bottom, n, mu, lpf)
    return list(mu[:n]), list(lpf[:n])


def _term(fn, *args):
    out = (U64 * 2)()
    fn(*args, out)
    return _s128(out[0], out[1])


def P2(x, y, threads=0):
    return _term(lib().orc_p2, *_split(x), y, threads)


def Ordinary(x, y, c=6, threads=0):
    return _term(lib().orc_ordinary, *_split(x), y, c, threads)


def S2_trivial(x, y, c=6):
    return _term(lib().orc_trivial, *_split(x), y, c)


def V(x, y, c=6, threads=0):
    return _term(lib().orc_easy, *_split(x), y, c, threads)


def V_a(x, y, c=6, threads=0):
    """V over x^(1/4) < p <= sqrt(x/y) (src/V.cu:58-87)."""
    return _term(lib().orc_easy_part, *_split(x), y, c, threads, 1)


def V_b(x, y, c=6, threads=0):
    """V over sqrt(x/y) < p <= cbrt(x) (src/V.cu:89-124)."""
    return _term(lib().orc_easy_part, *_split(x), y, c, threads, 2)


def S3(x, y, c=6, threads=0):
    return _term(lib().orc_hard, *_split(x), y, c, threads)


def sigma(x, y, threads=0):
    """The sigma partition of the easy leaves (oracle.c orc_sigma): dict of
    A, s1 .. s6 and S2 = A + s2 + s3 + s4 + 2 s5 + s6."""
    out = (U64 * 16)()
    lib().orc_sigma(*_split(x), y, threads, out)
    keys = ("A", "s1", "s2", "s3", "s4", "s5", "s6", "S2")
    return {k: _s128(out[2 * i], out[2 * i + 1]) for i, k in enumerate(keys)}


def pi_dr(x, y, c=6, threads=0):
    """Full DR evaluation -> dict of terms and pi(x)."""
    out = (U64 * 14)()
    lib().orc_pi_dr(*_split(x), y, c, threads, out)
    v = [_s128(out[2 * i], out[2 * i + 1]) for i in range(7)]
    keys = ("ordinary", "trivial", "easy", "hard", "p2", "pi_y", "pi")
    return dict(zip(keys, v))
