"""cudapix_amd -- MI355X combinatorial pi(x) (Deleglise-Rivat) engine.

Python mirror of the reference's pi(x) interface over the C ABI of
libcudapix_amd.so (include/cudapix.h).  Function names follow the reference
headers they replace: P2 (include/P2.cuh:16-17), Ordinary (include/ordinary.cuh:6-7),
S2_trivial (include/trivial.cuh:8-12), S3 (include/S3.cuh:3), V / V_a / V_b
(include/V.cuh:6-8), get_d_piTable / PiTable (include/pitable.cuh:7-51),
Phi (include/phi.cuh:17-36), gen_d_lpf / gen_d_mu (include/sieve/lpf_mu.cuh:14-21),
countPrimes / getDevicePrimes (CUDASieve), and pi_deleglise_rivat
(src/Deleglise-Rivat/deleglise-rivat.cpp:50).

Every computation runs in the HIP library on the GPU; there is no CPU
fallback.  Importing works without a GPU (the library loads and its host-side
self checks run); creating a Context needs a HIP device.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcudapix_amd.so")

U64 = ctypes.c_uint64
MASK64 = (1 << 64) - 1


class CpxU128(ctypes.Structure):
    _fields_ = [("lo", U64), ("hi", U64)]


CpxI128 = CpxU128


class CpxOpts(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("c", ctypes.c_uint32), ("alpha", ctypes.c_double),
                ("seg_threads", ctypes.c_uint32), ("seg_kwords", ctypes.c_uint32),
                ("hbm_budget", U64), ("fresh_tables", ctypes.c_uint32),
                ("use_graphs", ctypes.c_uint32), ("pass_bytes", U64)]


class CpxSigmaTerms(ctypes.Structure):
    _fields_ = [("a", CpxI128), ("sigma", CpxI128 * 6), ("s2", CpxI128), ("y", U64),
                ("exact", ctypes.c_uint32), ("pad", ctypes.c_uint32)]


class CpxTerms(ctypes.Structure):
    _fields_ = [("pi", CpxU128), ("ordinary", CpxI128), ("trivial", CpxI128),
                ("easy", CpxI128), ("hard", CpxI128), ("p2", CpxI128),
                ("y", U64), ("z", U64), ("c", U64), ("pi_y", U64),
                ("alpha", ctypes.c_double), ("ms", ctypes.c_double)]


class CpxBatchOpts(ctypes.Structure):
    _fields_ = [("use_graph", ctypes.c_uint32), ("tune", ctypes.c_uint32), ("p2_table", ctypes.c_uint32)]


class CpxBatchStats(ctypes.Structure):
    _fields_ = [("ms_total", ctypes.c_double), ("ms_plan", ctypes.c_double),
                ("ms_tables", ctypes.c_double), ("ms_device", ctypes.c_double),
                ("ms_segments", ctypes.c_double), ("tiles", U64), ("hard_segments", U64),
                ("p2_segments", U64), ("seg_threads", ctypes.c_uint32),
                ("seg_kwords", ctypes.c_uint32), ("passes", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


class CpxServerOpts(ctypes.Structure):
    _fields_ = [("hbm_budget", U64), ("max_batch", ctypes.c_uint32),
                ("max_wait_ms", ctypes.c_double), ("slots", ctypes.c_uint32), ("max_x", CpxU128),
                ("mode", ctypes.c_uint32)]


SERVER_STREAM = 0    # cohorts join one running streaming launch (continuous batching)
SERVER_BATCHES = 1   # each admitted FIFO prefix runs as a batch launch of its own


class CpxResult(ctypes.Structure):
    _fields_ = [("ticket", U64), ("x", CpxU128), ("pi", CpxU128), ("status", ctypes.c_int),
                ("t_arrive_ms", ctypes.c_double), ("t_admit_ms", ctypes.c_double),
                ("t_done_ms", ctypes.c_double), ("launch", U64), ("cohort", U64),
                ("dev_first_ms", ctypes.c_double), ("dev_last_ms", ctypes.c_double)]


class CpxServerStats(ctypes.Structure):
    _fields_ = [("submitted", U64), ("completed", U64), ("rejected", U64), ("cancelled", U64),
                ("batches", U64), ("max_inflight_bytes", U64), ("p50_ms", ctypes.c_double),
                ("p99_ms", ctypes.c_double), ("mean_ms", ctypes.c_double),
                ("qps", ctypes.c_double), ("launches", U64), ("joined", U64),
                ("arena_bytes", U64)]


# every symbol include/cudapix.h declares (checked by the CPU tests)
EXPORTS = [
    "cpx_opts_default", "cpx_ctx_create", "cpx_ctx_destroy", "cpx_strerror",
    "cpx_ctx_last_error", "cpx_choose_y", "cpx_ctx_device_bytes", "cpx_debug_fail_alloc",
    "cpx_pi_u64", "cpx_pi_u128",
    "cpx_p2_u64", "cpx_p2_u128", "cpx_div_x_by_y",
    "cpx_ordinary_u64", "cpx_ordinary_u128",
    "cpx_s2_trivial_u64", "cpx_s2_trivial_u128",
    "cpx_s3_u64", "cpx_s3_u128", "cpx_v_u64", "cpx_v_u128", "cpx_v_split_u64", "cpx_s2_sigma_u128",
    "cpx_count_primes", "cpx_get_device_primes", "cpx_get_d_pitable",
    "cpx_pitable_create", "cpx_pitable_destroy", "cpx_pitable_current",
    "cpx_pitable_next_up", "cpx_pitable_next_down", "cpx_pitable_info",
    "cpx_phi_create", "cpx_phi_destroy", "cpx_phi_eval", "cpx_phi_range",
    "cpx_gen_d_lpf", "cpx_gen_d_mu", "cpx_free", "cpx_gen_d_lpf64", "cpx_copy_to_host",
    "cpx_pitable_set_range", "cpx_pitable_set_pi_base", "cpx_pitable_max_range",
    "cpx_phi_create_pmax", "cpx_phi_lpf", "cpx_phi_prime_list", "cpx_phi_mod_cache",
    "cpx_pi_batch", "cpx_pi_batch_terms", "cpx_autotune",
    "cpx_server_create", "cpx_server_submit", "cpx_server_cancel", "cpx_server_poll",
    "cpx_server_drain", "cpx_server_stats_get", "cpx_server_destroy", "cpx_server_replay",
    "cpx_request_bytes", "cpx_selftest_arith", "cpx_selftest_wave", "cpx_probe_peaks",
    "cpx_get_tuning",
    "cpx_set_tuning",
    "cpx_pi_batch_submit", "cpx_pi_batch_collect", "cpx_pi_batch_noalloc",
    "cpx_get_device_primes32", "cpx_get_host_primes", "cpx_gen_device_bit_sieve",
]

ERRORS = {0: "CPX_OK", -1: "CPX_EINVAL", -2: "CPX_ENOMEM", -3: "CPX_EDEVICE", -4: "CPX_ERANGE",
          -5: "CPX_EAGAIN", -6: "CPX_ECANCELED", -7: "CPX_ENOTFOUND", -8: "CPX_ESTATE"}

_lib = None


class CpxError(RuntimeError):
    def __init__(self, code, msg=""):
        self.code = code
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "?"), code, msg))


def lib():
    """Load libcudapix_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libcudapix_amd.so is not built (run __graft_entry__.build()); "
                          "there is no fallback path")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    sig = {
        "cpx_opts_default": (None, [P(CpxOpts)]),
        "cpx_ctx_create": (ctypes.c_int, [P(CpxOpts), P(vp)]),
        "cpx_ctx_destroy": (ctypes.c_int, [vp]),
        "cpx_strerror": (ctypes.c_char_p, [ctypes.c_int]),
        "cpx_ctx_last_error": (ctypes.c_char_p, [vp]),
        "cpx_choose_y": (ctypes.c_int, [vp, CpxU128, P(U64)]),
        "cpx_ctx_device_bytes": (U64, [vp]),
        "cpx_debug_fail_alloc": (ctypes.c_int, [ctypes.c_int]),
        "cpx_pi_u64": (ctypes.c_int, [vp, U64, U64, ctypes.c_uint32, P(U64), P(CpxTerms)]),
        "cpx_pi_u128": (ctypes.c_int, [vp, CpxU128, U64, ctypes.c_uint32, P(CpxU128), P(CpxTerms)]),
        "cpx_p2_u64": (ctypes.c_int, [vp, U64, U64, P(U64)]),
        "cpx_p2_u128": (ctypes.c_int, [vp, CpxU128, U64, P(CpxU128)]),
        "cpx_div_x_by_y": (ctypes.c_int, [vp, CpxU128, vp, ctypes.c_size_t]),
        "cpx_ordinary_u64": (ctypes.c_int, [vp, U64, U64, ctypes.c_uint16, P(ctypes.c_int64)]),
        "cpx_ordinary_u128": (ctypes.c_int, [vp, CpxU128, U64, ctypes.c_uint16, P(CpxI128)]),
        "cpx_s2_trivial_u64": (ctypes.c_int, [vp, U64, U64, P(U64)]),
        "cpx_s2_trivial_u128": (ctypes.c_int, [vp, CpxU128, U64, P(CpxU128)]),
        "cpx_s3_u64": (ctypes.c_int, [vp, U64, U64, ctypes.c_uint32, P(ctypes.c_int64)]),
        "cpx_s3_u128": (ctypes.c_int, [vp, CpxU128, U64, ctypes.c_uint32, P(CpxI128)]),
        "cpx_v_u64": (ctypes.c_int, [vp, U64, U64, P(ctypes.c_int64)]),
        "cpx_v_u128": (ctypes.c_int, [vp, CpxU128, U64, P(CpxI128)]),
        "cpx_v_split_u64": (ctypes.c_int, [vp, U64, U64, P(ctypes.c_int64), P(ctypes.c_int64)]),
        "cpx_s2_sigma_u128": (ctypes.c_int, [vp, CpxU128, U64, P(CpxSigmaTerms)]),
        "cpx_count_primes": (ctypes.c_int, [vp, U64, U64, P(U64)]),
        "cpx_get_device_primes": (ctypes.c_int, [vp, U64, U64, P(vp), P(ctypes.c_size_t)]),
        "cpx_get_device_primes32": (ctypes.c_int, [vp, U64, U64, P(vp), P(ctypes.c_size_t)]),
        "cpx_get_host_primes": (ctypes.c_int, [vp, U64, U64, P(U64), ctypes.c_size_t,
                                               P(ctypes.c_size_t)]),
        "cpx_gen_device_bit_sieve": (ctypes.c_int, [vp, U64, U64, P(vp), P(ctypes.c_size_t)]),
        "cpx_get_d_pitable": (ctypes.c_int, [vp, ctypes.c_uint32, P(vp)]),
        "cpx_pitable_create": (ctypes.c_int, [vp, U64, U64, U64, P(vp)]),
        "cpx_pitable_destroy": (ctypes.c_int, [vp]),
        "cpx_pitable_current": (ctypes.c_int, [vp, P(vp)]),
        "cpx_pitable_next_up": (ctypes.c_int, [vp, P(vp)]),
        "cpx_pitable_next_down": (ctypes.c_int, [vp, P(vp)]),
        "cpx_pitable_info": (ctypes.c_int, [vp, P(U64), P(U64), P(U64)]),
        "cpx_phi_create": (ctypes.c_int, [vp, ctypes.c_uint32, P(vp)]),
        "cpx_phi_destroy": (ctypes.c_int, [vp]),
        "cpx_phi_eval": (ctypes.c_int, [vp, ctypes.c_uint32, ctypes.c_uint32, P(U64)]),
        "cpx_phi_range": (ctypes.c_int, [vp, ctypes.c_uint32, ctypes.c_uint32, P(vp)]),
        "cpx_gen_d_lpf": (ctypes.c_int, [vp, U64, U64, P(vp)]),
        "cpx_gen_d_lpf64": (ctypes.c_int, [vp, U64, U64, P(vp)]),
        "cpx_copy_to_host": (ctypes.c_int, [vp, vp, vp, ctypes.c_size_t]),
        "cpx_pitable_set_range": (ctypes.c_int, [vp, U64]),
        "cpx_pitable_set_pi_base": (ctypes.c_int, [vp, U64]),
        "cpx_pitable_max_range": (ctypes.c_int, [vp, P(U64)]),
        "cpx_phi_create_pmax": (ctypes.c_int, [vp, ctypes.c_uint32, ctypes.c_uint32, P(vp)]),
        "cpx_phi_lpf": (ctypes.c_int, [vp, P(vp), P(U64)]),
        "cpx_phi_prime_list": (ctypes.c_int, [vp, P(vp), P(vp), P(ctypes.c_uint32)]),
        "cpx_phi_mod_cache": (ctypes.c_int, [vp, ctypes.c_uint32, P(U64)]),
        "cpx_gen_d_mu": (ctypes.c_int, [vp, U64, U64, P(vp)]),
        "cpx_free": (ctypes.c_int, [vp, vp]),
        "cpx_pi_batch": (ctypes.c_int, [vp, P(CpxU128), ctypes.c_size_t, P(CpxU128),
                                        P(CpxBatchOpts), P(CpxBatchStats)]),
        "cpx_pi_batch_noalloc": (ctypes.c_int, [vp, P(CpxU128), ctypes.c_size_t, P(CpxU128)]),
        "cpx_pi_batch_terms": (ctypes.c_int, [vp, P(CpxU128), ctypes.c_size_t, P(CpxU128),
                                              P(CpxTerms), P(CpxBatchOpts), P(CpxBatchStats)]),
        "cpx_autotune": (ctypes.c_int, [vp, P(CpxU128), ctypes.c_size_t, P(ctypes.c_uint32),
                                        P(ctypes.c_uint32)]),
        "cpx_pi_batch_submit": (ctypes.c_int, [vp, P(CpxU128), ctypes.c_size_t, P(CpxBatchOpts),
                                               P(U64)]),
        "cpx_pi_batch_collect": (ctypes.c_int, [vp, U64, P(CpxU128), P(CpxBatchStats)]),
        "cpx_server_create": (ctypes.c_int, [vp, P(CpxServerOpts), P(vp)]),
        "cpx_server_submit": (ctypes.c_int, [vp, CpxU128, P(U64)]),
        "cpx_server_cancel": (ctypes.c_int, [vp, U64]),
        "cpx_server_poll": (ctypes.c_int, [vp, P(CpxResult), ctypes.c_size_t, P(ctypes.c_size_t)]),
        "cpx_server_drain": (ctypes.c_int, [vp]),
        "cpx_server_stats_get": (ctypes.c_int, [vp, P(CpxServerStats)]),
        "cpx_server_destroy": (ctypes.c_int, [vp]),
        "cpx_server_replay": (ctypes.c_int, [vp, P(CpxU128), P(ctypes.c_double), ctypes.c_size_t,
                                             P(U64)]),
        "cpx_request_bytes": (U64, [vp, CpxU128]),
        "cpx_selftest_arith": (ctypes.c_int, [U64, U64]),
        "cpx_selftest_wave": (ctypes.c_int, [vp, U64, ctypes.c_uint32, P(ctypes.c_uint32)]),
        "cpx_probe_peaks": (ctypes.c_int, [vp, P(ctypes.c_double), P(ctypes.c_double)]),
        "cpx_get_tuning": (ctypes.c_int, [vp, P(ctypes.c_uint32), P(ctypes.c_uint32),
                                          P(ctypes.c_double), P(ctypes.c_uint32)]),
        "cpx_set_tuning": (ctypes.c_int, [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double,
                                          ctypes.c_uint32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def u128(v):
    if v < 0 or v >> 128:
        raise ValueError("x must be a non-negative 128-bit integer")
    return CpxU128(v & MASK64, v >> 64)


def from_u128(s):
    return (s.hi << 64) | s.lo


def from_i128(s):
    v = (s.hi << 64) | s.lo
    return v - (1 << 128) if v >> 127 else v


def _terms_dict(t):
    return {"pi": from_u128(t.pi), "ordinary": from_i128(t.ordinary),
            "trivial": from_i128(t.trivial), "easy": from_i128(t.easy),
            "hard": from_i128(t.hard), "p2": from_i128(t.p2), "y": t.y, "z": t.z,
            "c": t.c, "pi_y": t.pi_y, "alpha": t.alpha, "ms": t.ms}


class Context:
    """A device context: resident tables + stream (one per thread of use)."""

    def __init__(self, device=0, c=6, alpha=0.0, seg_threads=0, seg_kwords=0, hbm_budget=0,
                 fresh_tables=False, use_graphs=True, pass_bytes=0):
        L = lib()
        o = CpxOpts()
        L.cpx_opts_default(ctypes.byref(o))
        o.device, o.c, o.alpha = device, c, alpha
        o.seg_threads, o.seg_kwords, o.hbm_budget = seg_threads, seg_kwords, hbm_budget
        o.fresh_tables, o.use_graphs = int(fresh_tables), int(use_graphs)
        o.pass_bytes = pass_bytes
        h = ctypes.c_void_p()
        rc = L.cpx_ctx_create(ctypes.byref(o), ctypes.byref(h))
        if rc:
            raise CpxError(rc, "cpx_ctx_create")
        self.h = h
        self._L = L
        self._inflight = {}            # ticket -> batch size of submitted batches

    def close(self):
        if self.h:
            self._L.cpx_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc:
            msg = self._L.cpx_ctx_last_error(self.h)
            raise CpxError(rc, "%s: %s" % (what, msg.decode() if msg else ""))

    # ---- full pi(x)
    def pi(self, x, y=0, c=0, terms=False):
        o = CpxU128()
        t = CpxTerms()
        self._chk(self._L.cpx_pi_u128(self.h, u128(x), y, c, ctypes.byref(o), ctypes.byref(t)),
                  "pi")
        return _terms_dict(t) if terms else from_u128(o)

    def choose_y(self, x):
        y = U64()
        self._chk(self._L.cpx_choose_y(self.h, u128(x), ctypes.byref(y)), "choose_y")
        return y.value

    # ---- terms
    def P2(self, x, y):
        o = CpxU128()
        self._chk(self._L.cpx_p2_u128(self.h, u128(x), y, ctypes.byref(o)), "P2")
        return from_u128(o)

    def Ordinary(self, x, y, c=6):
        o = CpxI128()
        self._chk(self._L.cpx_ordinary_u128(self.h, u128(x), y, c, ctypes.byref(o)), "Ordinary")
        return from_i128(o)

    def S2_trivial(self, x, y):
        o = CpxU128()
        self._chk(self._L.cpx_s2_trivial_u128(self.h, u128(x), y, ctypes.byref(o)), "S2_trivial")
        return from_u128(o)

    def S3(self, x, y, c=6):
        o = CpxI128()
        self._chk(self._L.cpx_s3_u128(self.h, u128(x), y, c, ctypes.byref(o)), "S3")
        return from_i128(o)

    def V(self, x, y):
        o = CpxI128()
        self._chk(self._L.cpx_v_u128(self.h, u128(x), y, ctypes.byref(o)), "V")
        return from_i128(o)

    def V_split(self, x, y):
        a, b = ctypes.c_int64(), ctypes.c_int64()
        self._chk(self._L.cpx_v_split_u64(self.h, x, y, ctypes.byref(a), ctypes.byref(b)),
                  "V_split")
        return a.value, b.value

    def s2_sigma(self, x, y=0):
        """The reference's second form of the easy leaves (S2.cu:30-70, sigma.cu):
        dict of A, s1 .. s6, S2 = A + s2 + s3 + s4 + 2 s5 + s6, the y used, and
        exact (S2 equals V: no prime in (x^(1/4), x/y^2])."""
        o = CpxSigmaTerms()
        self._chk(self._L.cpx_s2_sigma_u128(self.h, u128(x), y, ctypes.byref(o)), "s2_sigma")
        d = {"A": from_i128(o.a), "S2": from_i128(o.s2), "y": o.y, "exact": bool(o.exact)}
        for i in range(6):
            d["s%d" % (i + 1)] = from_i128(o.sigma[i])
        return d

    def V_a(self, x, y):
        """V over x^(1/4) < p <= sqrt(x/y) (include/V.cuh:7)."""
        return self.V_split(x, y)[0]

    def V_b(self, x, y):
        """V over sqrt(x/y) < p <= cbrt(x) (include/V.cuh:8)."""
        return self.V_split(x, y)[1]

    def terms_u64(self, x, y, c=6):
        """Every term through the 64-bit entry points, as the reference's
        64-bit callers bind them (include/P2.cuh:16, ordinary.cuh:6,
        trivial.cuh:8, S3.cuh:3, V.cuh:6, deleglise-rivat.cpp:50)."""
        L, h = self._L, self.h
        u, i = U64(), ctypes.c_int64()
        out = {}
        self._chk(L.cpx_p2_u64(h, x, y, ctypes.byref(u)), "P2 u64")
        out["p2"] = u.value
        self._chk(L.cpx_ordinary_u64(h, x, y, c, ctypes.byref(i)), "Ordinary u64")
        out["ordinary"] = i.value
        self._chk(L.cpx_s2_trivial_u64(h, x, y, ctypes.byref(u)), "S2_trivial u64")
        out["trivial"] = u.value
        self._chk(L.cpx_s3_u64(h, x, y, c, ctypes.byref(i)), "S3 u64")
        out["hard"] = i.value
        self._chk(L.cpx_v_u64(h, x, y, ctypes.byref(i)), "V u64")
        out["easy"] = i.value
        t = CpxTerms()
        self._chk(L.cpx_pi_u64(h, x, y, c, ctypes.byref(u), ctypes.byref(t)), "pi u64")
        out["pi"] = u.value
        return out

    # ---- primes / tables
    def countPrimes(self, lo, hi):
        o = U64()
        self._chk(self._L.cpx_count_primes(self.h, lo, hi, ctypes.byref(o)), "countPrimes")
        return o.value

    def _d2h(self, dptr, n, ctype):
        """Copy a device array into a Python list (through hipMemcpy in libamdhip64)."""
        buf = (ctype * max(n, 1))()
        hip = _hip()
        rc = hip.hipMemcpy(buf, dptr, ctypes.sizeof(ctype) * n, 2)  # DeviceToHost
        if rc:
            raise CpxError(-3, "hipMemcpy %d" % rc)
        return list(buf[:n])

    def getDevicePrimes(self, lo, hi):
        d = ctypes.c_void_p()
        n = ctypes.c_size_t()
        self._chk(self._L.cpx_get_device_primes(self.h, lo, hi, ctypes.byref(d), ctypes.byref(n)),
                  "getDevicePrimes")
        try:
            return self._d2h(d, n.value, U64)
        finally:
            self._L.cpx_free(self.h, d)

    def getDevicePrimes32(self, lo, hi):
        d = ctypes.c_void_p()
        n = ctypes.c_size_t()
        self._chk(self._L.cpx_get_device_primes32(self.h, lo, hi, ctypes.byref(d), ctypes.byref(n)),
                  "getDevicePrimes32")
        try:
            return self._d2h(d, n.value, ctypes.c_uint32)
        finally:
            self._L.cpx_free(self.h, d)

    def getHostPrimes(self, lo, hi):
        n = ctypes.c_size_t()
        self._chk(self._L.cpx_get_host_primes(self.h, lo, hi, None, 0, ctypes.byref(n)), "getHostPrimes")
        buf = (U64 * max(n.value, 1))()
        m = ctypes.c_size_t()
        self._chk(self._L.cpx_get_host_primes(self.h, lo, hi, buf, n.value, ctypes.byref(m)),
                  "getHostPrimes")
        return list(buf[:m.value])

    def genDeviceBitSieve(self, lo, hi):
        """Odd-only bit sieve words of [lo, hi) (lo a multiple of 64): bit i of
        word j is clear iff lo + 64 j + 2 i + 1 is an odd prime in [lo, hi)."""
        d = ctypes.c_void_p()
        n = ctypes.c_size_t()
        self._chk(self._L.cpx_gen_device_bit_sieve(self.h, lo, hi, ctypes.byref(d), ctypes.byref(n)),
                  "genDeviceBitSieve")
        try:
            return self._d2h(d, n.value, ctypes.c_uint32)
        finally:
            self._L.cpx_free(self.h, d)

    def get_d_piTable(self, hi):
        d = ctypes.c_void_p()
        self._chk(self._L.cpx_get_d_pitable(self.h, hi, ctypes.byref(d)), "get_d_piTable")
        try:
            return self._d2h(d, hi // 2, ctypes.c_uint32)
        finally:
            self._L.cpx_free(self.h, d)

    def gen_lpf(self, bottom, top):
        d = ctypes.c_void_p()
        self._chk(self._L.cpx_gen_d_lpf(self.h, bottom, top, ctypes.byref(d)), "gen_d_lpf")
        try:
            return self._d2h(d, (top - bottom) // 2, ctypes.c_uint32)
        finally:
            self._L.cpx_free(self.h, d)

    def gen_lpf64(self, bottom, top):
        """gen_d_lpf<uint64_t>: 64-bit entries, windows anywhere below 2^64 - 2^34."""
        d = ctypes.c_void_p()
        self._chk(self._L.cpx_gen_d_lpf64(self.h, bottom, top, ctypes.byref(d)), "gen_d_lpf64")
        try:
            return self._d2h(d, (top - bottom) // 2, U64)
        finally:
            self._L.cpx_free(self.h, d)

    def gen_mu(self, bottom, top):
        d = ctypes.c_void_p()
        self._chk(self._L.cpx_gen_d_mu(self.h, bottom, top, ctypes.byref(d)), "gen_d_mu")
        try:
            return self._d2h(d, (top - bottom) // 2, ctypes.c_int8)
        finally:
            self._L.cpx_free(self.h, d)

    def divXbyY(self, x, ys):
        hip = _hip()
        n = len(ys)
        arr = (U64 * max(n, 1))(*ys)
        d = ctypes.c_void_p()
        if hip.hipMalloc(ctypes.byref(d), 8 * max(n, 1)):
            raise CpxError(-2, "hipMalloc")
        try:
            hip.hipMemcpy(d, arr, 8 * n, 1)
            self._chk(self._L.cpx_div_x_by_y(self.h, u128(x), d, n), "divXbyY")
            return self._d2h(d, n, U64)
        finally:
            hip.hipFree(d)

    # ---- batch
    def pi_batch(self, xs, use_graph=True, tune=False, terms=False, stats=False, p2_table=0):
        """pi(x) for every x in one launch sequence.  p2_table: 0 = P2 by lookups
        in the resident pi table when >= 8 queries qualify (the table grows),
        1 = only where the resident table reaches, 2 = never."""
        n = len(xs)
        arr = (CpxU128 * max(n, 1))(*[u128(v) for v in xs])
        out = (CpxU128 * max(n, 1))()
        bo = CpxBatchOpts(int(use_graph), int(tune), int(p2_table))
        st = CpxBatchStats()
        if terms:
            tt = (CpxTerms * max(n, 1))()
            rc = self._L.cpx_pi_batch_terms(self.h, arr, n, out, tt, ctypes.byref(bo),
                                            ctypes.byref(st))
        else:
            rc = self._L.cpx_pi_batch(self.h, arr, n, out, ctypes.byref(bo), ctypes.byref(st))
        self._chk(rc, "pi_batch")
        res = [_terms_dict(tt[i]) for i in range(n)] if terms else \
            [from_u128(out[i]) for i in range(n)]
        if stats:
            return res, {f: getattr(st, f) for f, _ in CpxBatchStats._fields_}
        return res

    def pi_batch_noalloc(self, xs):
        """pi_batch with nothing allocated: the batch's state goes to the
        library's static 128 MiB device array; the resident tables must
        already cover it (CPX_EAGAIN), its state must fit (CPX_ENOMEM)."""
        n = len(xs)
        arr = (CpxU128 * max(n, 1))(*[u128(v) for v in xs])
        out = (CpxU128 * max(n, 1))()
        self._chk(self._L.cpx_pi_batch_noalloc(self.h, arr, n, out), "pi_batch_noalloc")
        return [from_u128(out[i]) for i in range(n)]

    def pi_batch_submit(self, xs, use_graph=True, p2_table=0):
        """Plan and queue a batch without waiting; returns a ticket for
        pi_batch_collect.  Two batches may be in flight per context: the next
        one is planned and uploaded while the previous one runs."""
        n = len(xs)
        arr = (CpxU128 * max(n, 1))(*[u128(v) for v in xs])
        bo = CpxBatchOpts(int(use_graph), 0, int(p2_table))
        t = U64()
        self._chk(self._L.cpx_pi_batch_submit(self.h, arr, n, ctypes.byref(bo), ctypes.byref(t)),
                  "pi_batch_submit")
        self._inflight[t.value] = n
        return t.value

    def pi_batch_collect(self, ticket, stats=False):
        """Wait for a submitted batch; its pi(x) values in submission order."""
        n = self._inflight.pop(ticket, None)     # collected (or failed) either way
        if n is None:
            raise CpxError(-7, "pi_batch_collect: unknown ticket")
        out = (CpxU128 * max(n, 1))()
        st = CpxBatchStats()
        self._chk(self._L.cpx_pi_batch_collect(self.h, U64(ticket), out, ctypes.byref(st)),
                  "pi_batch_collect")
        res = [from_u128(out[i]) for i in range(n)]
        if stats:
            return res, {f: getattr(st, f) for f, _ in CpxBatchStats._fields_}
        return res

    def pi_batches(self, batches, use_graph=True, p2_table=0):
        """pi(x) of every batch in an iterable of x lists, pipelined two deep:
        yields each batch's results in order."""
        pending = []
        try:
            for xs in batches:
                pending.append(self.pi_batch_submit(xs, use_graph=use_graph, p2_table=p2_table))
                if len(pending) == 2:
                    yield self.pi_batch_collect(pending.pop(0))
            while pending:
                yield self.pi_batch_collect(pending.pop(0))
        finally:
            # a consumer that stops early leaves batches in flight: collect
            # them so that the context's slots are free again
            for t in pending:
                try:
                    self.pi_batch_collect(t)
                except CpxError:
                    pass

    def autotune(self, xs):
        n = len(xs)
        arr = (CpxU128 * max(n, 1))(*[u128(v) for v in xs])
        t, k = ctypes.c_uint32(), ctypes.c_uint32()
        self._chk(self._L.cpx_autotune(self.h, arr, n, ctypes.byref(t), ctypes.byref(k)),
                  "autotune")
        return t.value, k.value

    def tuning(self):
        """(threads, kwords, alpha_scale, c) the context currently uses."""
        t, k, c = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        a = ctypes.c_double()
        self._chk(self._L.cpx_get_tuning(self.h, ctypes.byref(t), ctypes.byref(k), ctypes.byref(a),
                                         ctypes.byref(c)), "get_tuning")
        return {"seg_threads": t.value, "seg_kwords": k.value, "alpha_scale": a.value, "c": c.value}

    def set_tuning(self, seg_threads, seg_kwords, alpha_scale, c):
        """Use a tuning (as returned by tuning()) without re-timing it."""
        self._chk(self._L.cpx_set_tuning(self.h, int(seg_threads), int(seg_kwords), float(alpha_scale),
                                         int(c)), "set_tuning")

    def request_bytes(self, x):
        return self._L.cpx_request_bytes(self.h, u128(x))

    def device_bytes(self):
        return self._L.cpx_ctx_device_bytes(self.h)

    def probe_peaks(self):
        """Measured LDS read bandwidth (bytes/s) and int32 VALU rate (lane ops/s)."""
        a, b = ctypes.c_double(), ctypes.c_double()
        self._chk(self._L.cpx_probe_peaks(self.h, ctypes.byref(a), ctypes.byref(b)), "probe_peaks")
        return a.value, b.value

    def selftest_wave(self, seed=1, nwaves=4096):
        """Wavefronts whose register scan/sum differ from a serial scan."""
        bad = ctypes.c_uint32()
        self._chk(self._L.cpx_selftest_wave(self.h, seed, nwaves, ctypes.byref(bad)), "selftest_wave")
        return bad.value


class PiTable:
    """Sliding pi-table window (class PiTable, include/pitable.cuh:12-51)."""

    def __init__(self, ctx, base, bottom, range_):
        self.ctx = ctx
        h = ctypes.c_void_p()
        ctx._chk(ctx._L.cpx_pitable_create(ctx.h, base, bottom, range_, ctypes.byref(h)),
                 "PiTable")
        self.h = h

    def _get(self, fn):
        d = ctypes.c_void_p()
        self.ctx._chk(fn(self.h, ctypes.byref(d)), "PiTable window")
        base, rng, pb = U64(), U64(), U64()
        self.ctx._L.cpx_pitable_info(self.h, ctypes.byref(base), ctypes.byref(rng),
                                     ctypes.byref(pb))
        vals = self.ctx._d2h(d, rng.value // 2 + 1, ctypes.c_uint32)
        return base.value, pb.value, vals

    def getCurrent(self):
        return self._get(self.ctx._L.cpx_pitable_current)

    def getNextUp(self):
        return self._get(self.ctx._L.cpx_pitable_next_up)

    def getNextDown(self):
        return self._get(self.ctx._L.cpx_pitable_next_down)

    def set_pi_base(self, pi_base):
        self.ctx._chk(self.ctx._L.cpx_pitable_set_pi_base(self.h, pi_base), "PiTable.set_pi_base")

    def reallocate(self, range_):
        self.ctx._chk(self.ctx._L.cpx_pitable_set_range(self.h, range_), "PiTable.reallocate")

    def getMaxRange(self):
        r = U64()
        self.ctx._chk(self.ctx._L.cpx_pitable_max_range(self.ctx.h, ctypes.byref(r)), "PiTable.getMaxRange")
        return r.value

    def close(self):
        if self.h:
            self.ctx._L.cpx_pitable_destroy(self.h)
            self.h = None


class Phi:
    """phi(x, a) for x < max from a device LPF table (class Phi, include/phi.cuh:17-36)."""

    def __init__(self, ctx, max_, p_max=0):
        self.ctx = ctx
        h = ctypes.c_void_p()
        ctx._chk(ctx._L.cpx_phi_create_pmax(ctx.h, max_, p_max, ctypes.byref(h)), "Phi")
        self.h = h

    def get_d_lpf(self):
        """(device pointer, entries) of the object's odd-only LPF array."""
        d, n = ctypes.c_void_p(), U64()
        self.ctx._chk(self.ctx._L.cpx_phi_lpf(self.h, ctypes.byref(d), ctypes.byref(n)), "Phi.get_d_lpf")
        return d, n.value

    def lpf(self):
        d, n = self.get_d_lpf()
        return self.ctx._d2h(d, n, ctypes.c_uint32)

    def prime_list(self):
        """(host list, device list copied back) of get_h_primeList / get_d_primeList."""
        d, h, n = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_uint32()
        self.ctx._chk(self.ctx._L.cpx_phi_prime_list(self.h, ctypes.byref(d), ctypes.byref(h),
                                                     ctypes.byref(n)), "Phi.primeList")
        hv = list(ctypes.cast(h, ctypes.POINTER(ctypes.c_uint32))[:n.value]) if n.value else []
        return hv, self.ctx._d2h(d, n.value, ctypes.c_uint32)

    def makeModCache(self, pi_max):
        out = (U64 * max(pi_max - 1, 1))()
        self.ctx._chk(self.ctx._L.cpx_phi_mod_cache(self.h, pi_max, out), "Phi.makeModCache")
        return list(out[:max(pi_max - 1, 0)])

    def phi(self, x, a):
        o = U64()
        self.ctx._chk(self.ctx._L.cpx_phi_eval(self.h, x, a, ctypes.byref(o)), "Phi.phi")
        return o.value

    def generateRange(self, num, a):
        d = ctypes.c_void_p()
        self.ctx._chk(self.ctx._L.cpx_phi_range(self.h, num, a, ctypes.byref(d)),
                      "Phi.generateRange")
        try:
            return self.ctx._d2h(d, num // 2, ctypes.c_uint32)
        finally:
            self.ctx._L.cpx_free(self.ctx.h, d)

    def close(self):
        if self.h:
            self.ctx._L.cpx_phi_destroy(self.h)
            self.h = None


class Server:
    """pi(x) server.  mode SERVER_STREAM: continuous batching, requests join
    the running streaming launch as cohorts, admitted by the HBM bytes of
    their planned state; SERVER_BATCHES: each admitted FIFO prefix runs as a
    batch launch of its own while later arrivals queue."""

    def __init__(self, ctx, hbm_budget=0, max_batch=256, max_wait_ms=2.0, slots=4, max_x=0,
                 mode=SERVER_STREAM):
        self.ctx = ctx
        o = CpxServerOpts(hbm_budget, max_batch, max_wait_ms, slots, u128(max_x), mode)
        h = ctypes.c_void_p()
        ctx._chk(ctx._L.cpx_server_create(ctx.h, ctypes.byref(o), ctypes.byref(h)), "Server")
        self.h = h
        self._L = ctx._L

    def submit(self, x):
        t = U64()
        rc = self._L.cpx_server_submit(self.h, u128(x), ctypes.byref(t))
        if rc:
            raise CpxError(rc, "submit")
        return t.value

    def cancel(self, ticket):
        return self._L.cpx_server_cancel(self.h, ticket)

    def replay(self, xs, t_ms):
        n = len(xs)
        arr = (CpxU128 * max(n, 1))(*[u128(v) for v in xs])
        ts = (ctypes.c_double * max(n, 1))(*t_ms)
        tk = (U64 * max(n, 1))()
        rc = self._L.cpx_server_replay(self.h, arr, ts, n, tk)
        if rc:
            raise CpxError(rc, "replay")
        return list(tk[:n])

    def drain(self):
        self._L.cpx_server_drain(self.h)

    def poll(self, max_=4096):
        out = (CpxResult * max_)()
        n = ctypes.c_size_t()
        self._L.cpx_server_poll(self.h, out, max_, ctypes.byref(n))
        return [{"ticket": r.ticket, "x": from_u128(r.x), "pi": from_u128(r.pi),
                 "status": r.status, "t_arrive_ms": r.t_arrive_ms, "t_admit_ms": r.t_admit_ms,
                 "t_done_ms": r.t_done_ms, "launch": r.launch, "cohort": r.cohort,
                 "dev_first_ms": r.dev_first_ms, "dev_last_ms": r.dev_last_ms}
                for r in out[:n.value]]

    def stats(self):
        s = CpxServerStats()
        self._L.cpx_server_stats_get(self.h, ctypes.byref(s))
        return {f: getattr(s, f) for f, _ in CpxServerStats._fields_}

    def close(self):
        if self.h:
            self._L.cpx_server_destroy(self.h)
            self.h = None


_hip_lib = None


def _hip():
    global _hip_lib
    if _hip_lib is None:
        h = ctypes.CDLL("libamdhip64.so")
        h.hipMemcpy.restype = ctypes.c_int
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipMalloc.restype = ctypes.c_int
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.restype = ctypes.c_int
        h.hipFree.argtypes = [ctypes.c_void_p]
        _hip_lib = h
    return _hip_lib


_default = None


def debug_fail_alloc(n):
    """Test hook: the n-th device allocation of the library from now on fails
    as on an exhausted device (0 disarms)."""
    rc = lib().cpx_debug_fail_alloc(int(n))
    if rc:
        raise CpxError(rc, "debug_fail_alloc")


def default_context():
    global _default
    if _default is None:
        _default = Context()
    return _default


# reference-named module-level functions on a default context
def P2(x, y):
    return default_context().P2(x, y)


def Ordinary(x, y, c=6):
    return default_context().Ordinary(x, y, c)


def S2_trivial(x, y):
    return default_context().S2_trivial(x, y)


def S3(x, y, c=6):
    return default_context().S3(x, y, c)


def V(x, y):
    return default_context().V(x, y)


def pi_deleglise_rivat(x, y=0, c=6):
    return default_context().pi(x, y, c)


def selftest_arith(seed=1, n=100000):
    return lib().cpx_selftest_arith(seed, n)
